"""ctypes binding of libso3ft.so (declared in include/so3ft.h).

The library is loaded from this package directory.  If it is missing the import
of the transform API fails loudly: there is no CPU fallback.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libso3ft.so")
HEADER = os.path.join(HERE, "..", "include", "so3ft.h")

_i64, _i32, _int, _dbl = C.c_int64, C.c_int32, C.c_int, C.c_double
_vp = C.c_void_p
_P = C.POINTER

SIGNATURES = {
    "so3ft_get_kostelec_recurrence_val": (_int, []),
    "so3ft_get_risbo_recurrence_val": (_int, []),
    "so3ft_last_error": (C.c_char_p, []),
    "so3ft_init": (_int, [_i64, _i64, _int, _int, _i64, _i32, _P(_i64)]),
    "so3ft_destroy": (_int, [_i64]),
    "so3ft_reset": (_int, [_i64, _i64, _i64, _int, _int, _i64, _i32]),
    "so3ft_init_fft": (_int, [_i64, _int]),
    "so3ft_get_bw": (_int, [_i64, _P(_i64)]),
    "so3ft_get_lmax": (_int, [_i64, _P(_i64)]),
    "so3ft_set_lmax": (_int, [_i64, _i64]),
    "so3ft_get_recurrence_type": (_int, [_i64, _P(_i64)]),
    "so3ft_wigners_are_precomputed": (_int, [_i64, _P(_int)]),
    "so3ft_get_fftw_flags": (_int, [_i64, _P(_i32)]),
    "so3ft_set_dwt_engine": (_int, [_i64, _int]),
    "so3ft_get_dwt_engine": (_int, [_i64, _P(_int)]),
    "so3ft_isoft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_soft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_isoft_many": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_soft_many": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_irsoft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_rsoft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_irsoft_many": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_rsoft_many": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_isoft_many_c64": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_soft_many_c64": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_soft_many_c64_noalloc": (_int, [_i64, _vp, _vp, _i64, _vp]),
    "so3ft_inverse_wigner_trf_cmplx": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_forward_wigner_trf_cmplx": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_fft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_ifft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_rfft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_irfft": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_integrate_over_so3_cmplx": (_int, [_i64, _vp, _P(_dbl), _P(_dbl), _vp]),
    "so3ft_integrate_over_so3_real": (_int, [_i64, _vp, _P(_dbl), _vp]),
    "so3ft_cross_correlation_ylm_cmplx": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "so3ft_cross_correlation_ylm_cmplx_many": (_int, [_i64, _vp, _vp, _vp, _i64, _vp]),
    "so3ft_cross_correlation_argmax_many": (_int, [_i64, _vp, _vp, _i64, _vp, _vp, _vp]),
    "so3ft_cross_correlation_ylm_cmplx_3d": (_int, [_i64, _vp, _vp, _vp, _P(_dbl), _i64, _i64, _i64, _vp]),
    "so3ft_cross_correlation_ylm_real": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "so3ft_cross_correlation_ylm_real_3d": (_int, [_i64, _vp, _vp, _vp, _P(_dbl), _i64, _i64, _i64, _vp]),
    "so3ft_inverse_wigner_trf_corr_cmplx": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "so3ft_inverse_wigner_trf_corr_real": (_int, [_i64, _vp, _vp, _vp, _vp]),
    "so3ft_wigner_mn_recurrence_risbo": (_int, [_vp, _i64, _vp, _vp, _vp, _i64, _i64, _vp, _vp]),
    "so3ft_wigner_recurrence_risbo_reduced": (_int, [_vp, _i64, _vp, _vp, _vp, _i64, _i64, _int, _vp, _vp]),
    "so3ft_genwig_all": (_int, [_i64, _i64, _vp, _vp]),
    "so3ft_full_big_d_risbo": (_int, [_i64, _dbl, _dbl, _dbl, _int, _vp, _vp]),
    "so3ft_enforce_real_sym_mnl": (_int, [_i64, _vp, _i64, _vp]),
    "so3ft_enforce_real_sym_lmn": (_int, [_i64, _vp, _i64, _vp]),
    "so3ft_rotate_ylm_cmplx": (_int, [_i64, _vp, _i64, _vp, _i64, _int, _vp, _vp]),
    "so3ft_mlp_last_error": (C.c_char_p, []),
    "so3ft_mlp_create": (_int, [_P(_i32), _int, _P(_i64)]),
    "so3ft_mlp_destroy": (_int, [_i64]),
    "so3ft_mlp_num_params": (_int, [_i64, _P(_i64)]),
    "so3ft_mlp_set_params": (_int, [_i64, _vp]),
    "so3ft_mlp_get_params": (_int, [_i64, _int, _vp]),
    "so3ft_mlp_forward": (_int, [_i64, _vp, _i64, _vp, _vp, _vp]),
    "so3ft_mlp_train_step": (_int, [_i64, _vp, _vp, _i64, C.c_float, _P(C.c_float), _vp]),
    "so3ft_align_features": (_int, [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _vp]),
    "so3ft_gemm_bf16": (_int, [_vp, _vp, _vp, _int, _int, _int, _vp]),
    "so3ft_refine_correlation_peaks": (_int, [_i64, _vp, _vp, _i64, _vp, _vp, _vp, _i64, _i64, _vp]),
    "so3ft_group_layout": (_int, [_vp, _i64, _vp, _vp, _P(_i64), _P(_i64)]),
    "so3ft_group_create": (_int, [_vp, _i64, _P(_i64)]),
    "so3ft_group_destroy": (_int, [_i64]),
    "so3ft_group_isoft_c64": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_group_soft_c64": (_int, [_i64, _vp, _vp, _vp]),
    "so3ft_group_autotune": (_int, [_i64, _vp]),
    "so3ft_group_get_config": (_int, [_i64, _vp]),
    "so3ft_group_set_config": (_int, [_i64, _vp]),
    "so3ft_group_set_graphs": (_int, [_i64, _int]),
    "so3ft_bench_kernel": (_int, [_i64, _int, _int, _vp, _vp, _i64, _vp, _P(C.c_float)]),
}

_lib = None


class So3ftError(RuntimeError):
    pass


def header_symbols(path=HEADER):
    """Function names declared in include/so3ft.h."""
    with open(path) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(so3ft_[a-z0-9_]+)\s*\(", txt)))


def lib():
    """Load libso3ft.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise So3ftError(
                f"{LIB_PATH} not found: build it with `python -m pysofft_amd.build` "
                "(there is no CPU fallback)")
        h = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(rc):
    if rc != 0:
        msg = lib().so3ft_last_error()
        raise So3ftError(f"so3ft error {rc}: {msg.decode() if msg else ''}")
    return rc
