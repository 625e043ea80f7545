"""Drop-in ``Soft`` transform class (API of pysofft/soft.py) on top of libso3ft.

Every transform runs on the GPU through the C ABI of include/so3ft.h.  Inputs may
be numpy arrays (copied to the device and the result copied back into ``out``, as
the reference returns numpy arrays) or torch CUDA tensors (no copies; results are
torch tensors on the same device and stream).  There is no CPU fallback: without
the built library or a GPU the calls raise.

Array layouts are the reference's: coefficients ``((4bw^3-bw)/3,)`` complex in
"mnl" order, functions ``(2bw, 2bw, 2bw)`` C-order (gamma, alpha, beta).
"""
import ctypes as C
import multiprocessing
import os
from collections import namedtuple

import numpy as np

from . import _lib
from . import utils
from .wigner import create_alpha_gamma_samples, create_beta_samples


def log(txt):
    print(txt)


def _torch():
    import torch
    return torch


def _is_torch(x):
    try:
        import torch
    except ImportError:  # pragma: no cover
        return False
    return isinstance(x, torch.Tensor)


def _stream_ptr(dev):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)


class _Dev:
    """Device staging of one argument: torch tensors pass through, numpy is copied."""

    def __init__(self, x, dtype, device, shape=None):
        torch = _torch()
        self.host = None
        if _is_torch(x):
            if not x.is_cuda:
                x = x.to(device)
            if x.dtype != dtype:
                x = x.to(dtype)
            self.t = x.contiguous()
        else:
            a = np.asarray(x)
            nd = {torch.complex128: np.complex128, torch.float64: np.float64,
                  torch.complex64: np.complex64, torch.float32: np.float32}[dtype]
            a = np.ascontiguousarray(a, dtype=nd)
            self.t = torch.from_numpy(a).to(device)
        if shape is not None and tuple(self.t.shape[-len(shape):]) != tuple(shape):
            raise ValueError(f"expected trailing shape {shape}, got {tuple(self.t.shape)}")

    @property
    def ptr(self):
        return C.c_void_p(self.t.data_ptr())


class Soft:
    r"""SO(3) Fourier transforms of bandwidth ``bw`` (pysofft/soft.py:15-807)."""

    _handle = None
    recurrence_types = namedtuple('RecurrenceTypes', ['kostelec', 'risbo'])(0, 1)

    def __init__(self, bw: int, lmax: int | None = None, precompute_wigners: bool = False,
                 recurrence_type: int | None = None, init_ffts: bool = False, fftw_flags: int = 0,
                 use_fftw_wisdom: bool = False, device=None):
        torch = _torch()
        self._init_process_name = multiprocessing.current_process().name
        if recurrence_type not in self.recurrence_types:
            recurrence_type = self.recurrence_types.kostelec
        if lmax is None:
            lmax = bw - 1
        if not torch.cuda.is_available():
            raise _lib.So3ftError("pysofft_amd needs a ROCm GPU (torch.cuda.is_available() is False)")
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.use_fftw_wisdom = use_fftw_wisdom
        L = _lib.lib()
        h = C.c_int64()
        with torch.cuda.device(self.device):
            _lib.check(L.so3ft_init(int(bw), int(lmax), int(bool(precompute_wigners)), int(bool(init_ffts)),
                                    int(recurrence_type), int(fftw_flags), C.byref(h)))
        self._handle = h.value
        self._lmns = None

    def __del__(self):
        # freed only by the creating process (soft.py:69-72); at interpreter shutdown the
        # modules this needs may already be gone
        try:
            if self._handle is not None and multiprocessing.current_process().name == self._init_process_name:
                _lib.lib().so3ft_destroy(self._handle)
        except Exception:
            pass
        self._handle = None

    # -- properties (pysofft/soft.py:74-112) -------------------------------------
    def _get(self, fn, ctype=C.c_int64):
        v = ctype()
        _lib.check(fn(self._handle, C.byref(v)))
        return v.value

    @property
    def bw(self):
        return self._get(_lib.lib().so3ft_get_bw)

    @property
    def lmax(self):
        return self._get(_lib.lib().so3ft_get_lmax)

    @lmax.setter
    def lmax(self, value):
        if isinstance(value, (int, np.integer, np.ndarray)):
            _lib.check(_lib.lib().so3ft_set_lmax(self._handle, int(value)))
        else:
            log(f'Warning: lmax ({value}) not an integer, abbort changing lmax.')

    @property
    def recurrence_type(self):
        return self._get(_lib.lib().so3ft_get_recurrence_type)

    @property
    def wigners_are_precomputed(self):
        return bool(self._get(_lib.lib().so3ft_wigners_are_precomputed, C.c_int))

    @property
    def fftw_flags(self):
        return self._get(_lib.lib().so3ft_get_fftw_flags, C.c_int32)

    @property
    def dwt_engine(self):
        """FP64 DWT engine: "valu" (default: inverse on the FP64 vector ALUs, forward on
        FP64 MFMA with lane-beta recurrence chains) or "mfma" (both directions on
        v_mfma_f64_16x16x4f64 with per-pair d tiles)."""
        return {0: "valu", 1: "mfma"}[self._get(_lib.lib().so3ft_get_dwt_engine, C.c_int)]

    @dwt_engine.setter
    def dwt_engine(self, value):
        code = {"valu": 0, "mfma": 1}.get(value)
        if code is None:
            raise ValueError("dwt_engine must be 'valu' or 'mfma'")
        _lib.check(_lib.lib().so3ft_set_dwt_engine(self._handle, code))

    @property
    def coeff_order(self):
        """'lmn' for Risbo handles without precomputed Wigners, else 'mnl' (soft.py:98-106)."""
        if self.recurrence_type == self.recurrence_types.risbo and not self.wigners_are_precomputed:
            return 'lmn'
        return 'mnl'

    @property
    def coeff_indices(self):
        if self._lmns is None:
            if self.coeff_order == 'mnl':
                self._lmns = utils.get_coeff_degrees(self.bw)
            else:
                self._lmns = utils.get_coeff_degrees_risbo(self.bw)
            self._lmns.flags.writeable = False
        return self._lmns

    @property
    def euler_angles(self):
        beta = create_beta_samples(self.bw * 2)
        alpha = create_alpha_gamma_samples(self.bw * 2)
        return {'gamma': alpha.copy(), 'alpha': alpha, 'beta': beta}

    @property
    def n_coeffs(self):
        return utils.total_num_coeffs(self.bw)

    @property
    def grid_shape(self):
        return (2 * self.bw,) * 3

    # -- array factories (pysofft/soft.py:114-253) -------------------------------
    @staticmethod
    def _fill_random(arr, seed=12345):
        rng = np.random.default_rng(seed)
        arr[...] = rng.random(arr.shape)
        if np.iscomplexobj(arr):
            arr += 1.j * rng.random(arr.shape)
        return arr

    def get_coeff(self, real: bool = False, random: bool = False, seed: int = 12345, raw: bool = False,
                  howmany: int = 0):
        n_coeff = utils.total_num_coeffs(self.bw)
        shape = (howmany, n_coeff) if howmany > 0 else (n_coeff,)
        coeff = np.zeros(shape, dtype=np.complex128)
        lmn = self.coeff_order == 'lmn'
        if random:
            coeff = self._fill_random(coeff, seed=seed)
            if real:
                if howmany > 0:
                    (utils.enforce_real_sym_lmn_many if lmn else utils.enforce_real_sym_mnl_many)(coeff, self.bw)
                else:
                    (utils.enforce_real_sym_lmn if lmn else utils.enforce_real_sym_mnl)(coeff, self.bw)
        if not raw:
            coeff = CoeffSO3(coeff, self.coeff_indices, coeff_order=self.coeff_order)
        return coeff

    def get_so3func(self, real=False, random=False, seed=12345, howmany=0):
        shape = tuple(utils.euler_shape(self.bw).tolist())
        if howmany > 0:
            shape = (howmany,) + shape
        func = np.zeros(shape, dtype=np.float64 if real else np.complex128)
        if random:
            func = self._fill_random(func, seed=seed)
        return func

    def enforce_real_symmetry(self, coeff):
        lmn = self.coeff_order == 'lmn'
        if _is_torch(coeff):
            t = coeff
            n = 1 if t.ndim == 1 else t.shape[0]
            fn = _lib.lib().so3ft_enforce_real_sym_lmn if lmn else _lib.lib().so3ft_enforce_real_sym_mnl
            _lib.check(fn(self.bw, C.c_void_p(t.data_ptr()), n, _stream_ptr(t.device)))
            return coeff
        (utils.enforce_real_sym_lmn if lmn else utils.enforce_real_sym_mnl)(coeff, self.bw)

    # -- transforms (pysofft/soft.py:255-517) ------------------------------------
    def _partial_coeffs(self):
        """True when a forward transform writes only part of the coefficients: the
        Kostelec loops stop at max(|m|,|n|) <= lmax (soft.f08:1024-1065) and the Risbo
        loops at l <= lmax (soft.f08:1155-1183); the rest keep the values of ``out``,
        zeros for a fresh array (soft.py:289-294 allocates it with get_coeff)."""
        return self.lmax < self.bw - 1

    def _run(self, fn, x, xdtype, xshape, out, odtype, oshape, many=False, partial=False):
        torch = _torch()
        with torch.cuda.device(self.device):
            xin = _Dev(x, xdtype, self.device, xshape)
            n = xin.t.shape[0] if many else 1
            if many and xin.t.ndim != len(xshape) + 1:
                raise ValueError(f"expected (n,)+{xshape}, got {tuple(xin.t.shape)}")
            full = ((n,) + oshape) if many else oshape
            dev_out = out is not None and _is_torch(out) and out.is_cuda and out.dtype == odtype \
                and out.is_contiguous() and tuple(out.shape) == full
            if dev_out:
                o = out
            elif partial and out is not None:
                o = _Dev(out, odtype, self.device).t.reshape(full).clone()
            elif partial:
                o = torch.zeros(full, dtype=odtype, device=self.device)
            else:
                o = torch.empty(full, dtype=odtype, device=self.device)
            s = _stream_ptr(self.device)
            if many:
                _lib.check(fn(self._handle, xin.ptr, C.c_void_p(o.data_ptr()), n, s))
            else:
                _lib.check(fn(self._handle, xin.ptr, C.c_void_p(o.data_ptr()), s))
            if _is_torch(x):
                if out is not None and not dev_out:
                    out.copy_(o)
                    return out
                return o
            res = o.cpu().numpy()
            if out is None:
                return res
            np.copyto(out, res.reshape(np.shape(out)) if np.shape(out) != res.shape else res)
            return out

    def _coeff_out(self, out, howmany=0, real=False):
        return out

    def soft(self, so3func, out=None, use_mp: bool = False):
        torch = _torch()
        L = _lib.lib()
        res = self._run(L.so3ft_soft, so3func, torch.complex128, self.grid_shape, out, torch.complex128,
                        (self.n_coeffs,), partial=self._partial_coeffs())
        return self._wrap(res, out)

    def isoft(self, coeff, out=None, use_mp: bool = False):
        torch = _torch()
        L = _lib.lib()
        return self._run(L.so3ft_isoft, coeff, torch.complex128, (self.n_coeffs,), out, torch.complex128,
                         self.grid_shape)

    def rsoft(self, so3func, out=None, use_mp: bool = False):
        torch = _torch()
        L = _lib.lib()
        res = self._run(L.so3ft_rsoft, so3func, torch.float64, self.grid_shape, out, torch.complex128,
                        (self.n_coeffs,), partial=self._partial_coeffs())
        return self._wrap(res, out)

    def irsoft(self, coeff, out=None, use_mp: bool = False):
        torch = _torch()
        L = _lib.lib()
        return self._run(L.so3ft_irsoft, coeff, torch.complex128, (self.n_coeffs,), out, torch.float64,
                         self.grid_shape)

    def soft_many(self, so3funcs, out=None, use_mp: bool = False):
        torch = _torch()
        res = self._run(_lib.lib().so3ft_soft_many, so3funcs, torch.complex128, self.grid_shape, out,
                        torch.complex128, (self.n_coeffs,), many=True, partial=self._partial_coeffs())
        return self._wrap(res, out)

    def isoft_many(self, coeffs, out=None, use_mp=False):
        torch = _torch()
        return self._run(_lib.lib().so3ft_isoft_many, coeffs, torch.complex128, (self.n_coeffs,), out,
                         torch.complex128, self.grid_shape, many=True)

    def rsoft_many(self, so3funcs, out=None, use_mp: bool = False):
        torch = _torch()
        res = self._run(_lib.lib().so3ft_rsoft_many, so3funcs, torch.float64, self.grid_shape, out,
                        torch.complex128, (self.n_coeffs,), many=True, partial=self._partial_coeffs())
        return self._wrap(res, out)

    def irsoft_many(self, coeffs, out=None, use_mp: bool = False):
        torch = _torch()
        return self._run(_lib.lib().so3ft_irsoft_many, coeffs, torch.complex128, (self.n_coeffs,), out,
                         torch.float64, self.grid_shape, many=True)

    # -- mixed precision (no counterpart in the reference): complex64 storage, FP64
    # recurrence and FFT butterflies, FP32 MFMA contraction with hi/lo-split d
    def isoft_many_c64(self, coeffs, out=None):
        torch = _torch()
        return self._run(_lib.lib().so3ft_isoft_many_c64, coeffs, torch.complex64, (self.n_coeffs,), out,
                         torch.complex64, self.grid_shape, many=True)

    def soft_many_c64(self, so3funcs, out=None):
        torch = _torch()
        return self._run(_lib.lib().so3ft_soft_many_c64, so3funcs, torch.complex64, self.grid_shape, out,
                         torch.complex64, (self.n_coeffs,), many=True, partial=self._partial_coeffs())

    def _wrap(self, res, out):
        if out is None and isinstance(res, np.ndarray):
            return CoeffSO3(res, self.coeff_indices, coeff_order=self.coeff_order)
        return res

    # stage-level entry points (pysofft/soft.py:255-262)
    def _inverse_wigner_trf_cmplx(self, coeff, so3func=None, use_mp=False):
        torch = _torch()
        return self._run(_lib.lib().so3ft_inverse_wigner_trf_cmplx, coeff, torch.complex128, (self.n_coeffs,),
                         so3func, torch.complex128, self.grid_shape)

    def _forward_wigner_trf_cmplx(self, so3func, coeff=None, use_mp=False):
        torch = _torch()
        return self._run(_lib.lib().so3ft_forward_wigner_trf_cmplx, so3func, torch.complex128, self.grid_shape,
                         coeff, torch.complex128, (self.n_coeffs,), partial=self._partial_coeffs())

    def _fft(self, f1, f2=None):
        torch = _torch()
        return self._run(_lib.lib().so3ft_fft, f1, torch.complex128, self.grid_shape, f2, torch.complex128,
                         self.grid_shape)

    def _ifft(self, f1, f2=None):
        torch = _torch()
        return self._run(_lib.lib().so3ft_ifft, f1, torch.complex128, self.grid_shape, f2, torch.complex128,
                         self.grid_shape)

    def _rfft(self, f1, f2=None):
        """soft.py:708-709: real (2bw)^3 -> half-complex (2bw, bw+1, 2bw), /(2bw)."""
        torch = _torch()
        half = (2 * self.bw, self.bw + 1, 2 * self.bw)
        return self._run(_lib.lib().so3ft_rfft, f1, torch.float64, self.grid_shape, f2, torch.complex128, half)

    def _irfft(self, f1, f2=None):
        """soft.py:710-711: half-complex (2bw, bw+1, 2bw) -> real (2bw)^3, /(2bw)."""
        torch = _torch()
        half = (2 * self.bw, self.bw + 1, 2 * self.bw)
        return self._run(_lib.lib().so3ft_irfft, f1, torch.complex128, half, f2, torch.float64, self.grid_shape)

    # -- derived (pysofft/soft.py:519-703) ---------------------------------------
    def integrate_over_so3(self, f):
        torch = _torch()
        is_real = not (np.iscomplexobj(f) if not _is_torch(f) else f.is_complex())
        with torch.cuda.device(self.device):
            x = _Dev(f, torch.float64 if is_real else torch.complex128, self.device, self.grid_shape)
            s = _stream_ptr(self.device)
            re, im = C.c_double(), C.c_double()
            if is_real:
                _lib.check(_lib.lib().so3ft_integrate_over_so3_real(self._handle, x.ptr, C.byref(re), s))
                return float(re.value)
            _lib.check(_lib.lib().so3ft_integrate_over_so3_cmplx(self._handle, x.ptr, C.byref(re), C.byref(im), s))
            return complex(re.value, im.value)

    def cross_correlation_ylm_cmplx(self, f_lm, g_lm, out=None, use_mp: bool = False):
        torch = _torch()
        nb = self.bw ** 2
        with torch.cuda.device(self.device):
            f = _Dev(f_lm, torch.complex128, self.device, (nb,))
            g = _Dev(g_lm, torch.complex128, self.device, (nb,))
            o = torch.empty(self.grid_shape, dtype=torch.complex128, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_ylm_cmplx(self._handle, f.ptr, g.ptr,
                                                                   C.c_void_p(o.data_ptr()),
                                                                   _stream_ptr(self.device)))
            return self._finish(o, f_lm, out)

    # -- batched correlation (no counterpart in the reference, whose callers loop) ----
    def cross_correlation_ylm_cmplx_many(self, f_lms, g_lms, out=None):
        """(n, bw^2) LMc pairs -> (n, 2bw, 2bw, 2bw) correlation grids."""
        torch = _torch()
        nb = self.bw ** 2
        with torch.cuda.device(self.device):
            f = _Dev(f_lms, torch.complex128, self.device, None)
            g = _Dev(g_lms, torch.complex128, self.device, None)
            if f.t.ndim != 2 or f.t.shape[1] != nb or tuple(g.t.shape) != tuple(f.t.shape):
                raise ValueError(f"expected two (n, {nb}) arrays")
            n = f.t.shape[0]
            o = torch.empty((n,) + self.grid_shape, dtype=torch.complex128, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_ylm_cmplx_many(
                self._handle, f.ptr, g.ptr, C.c_void_p(o.data_ptr()), n, _stream_ptr(self.device)))
            return self._finish(o, f_lms, out)

    def correlation_peaks(self, f_lms, g_lms):
        """Correlate (n, bw^2) pairs and locate each peak of |C| without keeping the grids.

        Returns (index, value): index (n, 3) int64 grid indices (i_alpha, i_beta, i_gamma)
        of the peak, i.e. the rotation (alpha, beta, gamma) = euler_angles at those
        indices, and value (n,) = max |C|.  Arrays are torch CUDA tensors for torch
        inputs, numpy otherwise."""
        torch = _torch()
        nb = self.bw ** 2
        N = 2 * self.bw
        with torch.cuda.device(self.device):
            f = _Dev(f_lms, torch.complex128, self.device, None)
            g = _Dev(g_lms, torch.complex128, self.device, None)
            if f.t.ndim != 2 or f.t.shape[1] != nb or tuple(g.t.shape) != tuple(f.t.shape):
                raise ValueError(f"expected two (n, {nb}) arrays")
            n = f.t.shape[0]
            idx = torch.empty(n, dtype=torch.int64, device=self.device)
            val = torch.empty(n, dtype=torch.float64, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_argmax_many(
                self._handle, f.ptr, g.ptr, n, C.c_void_p(idx.data_ptr()), C.c_void_p(val.data_ptr()),
                _stream_ptr(self.device)))
            # flat C-order (gamma, alpha, beta) -> (i_alpha, i_beta, i_gamma)
            ig, rem = idx // (N * N), idx % (N * N)
            ia, ib = rem // N, rem % N
            out = torch.stack([ia, ib, ig], dim=1)
            if _is_torch(f_lms):
                return out, val
            return out.cpu().numpy(), val.cpu().numpy()

    def cross_correlation_ylm_cmplx_3d(self, f_lms, g_lms, out=None, radial_sampling_points=None,
                                       radial_limits=None, use_mp=False):
        torch = _torch()
        nb = self.bw ** 2
        n = len(f_lms)
        if radial_sampling_points is None:
            radial_sampling_points = np.linspace(0, 1, n)
        if radial_limits is None:
            radial_limits = [0, n]
        rad = np.ascontiguousarray(radial_sampling_points, dtype=np.float64)
        with torch.cuda.device(self.device):
            f = _Dev(f_lms, torch.complex128, self.device, (nb,))
            g = _Dev(g_lms, torch.complex128, self.device, (nb,))
            o = torch.empty(self.grid_shape, dtype=torch.complex128, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_ylm_cmplx_3d(
                self._handle, f.ptr, g.ptr, C.c_void_p(o.data_ptr()),
                rad.ctypes.data_as(C.POINTER(C.c_double)), len(rad), int(radial_limits[0]),
                int(radial_limits[1]) - 1, _stream_ptr(self.device)))
            return self._finish(o, f_lms, out)

    def _mlr_sets(self, x, many):
        """Complex128 MLr sets on the device: (bw(bw+1)/2,) (or (n, .)); longer trailing
        axes (the reference documents (bw**2,)) are cut to the used entries."""
        torch = _torch()
        nr = self.bw * (self.bw + 1) // 2
        d = _Dev(x, torch.complex128, self.device, None)
        t = d.t
        if t.ndim != (2 if many else 1) or t.shape[-1] < nr:
            raise ValueError(f"expected MLr coefficient sets with >= {nr} entries, got shape {tuple(t.shape)}")
        d.t = t[..., :nr].contiguous()
        return d

    def cross_correlation_ylm_real(self, f_lm, g_lm, out=None, use_mp=False):
        """Correlation of two real functions from MLr coefficients (pysofft/soft.py:594-626)."""
        torch = _torch()
        with torch.cuda.device(self.device):
            f = self._mlr_sets(f_lm, False)
            g = self._mlr_sets(g_lm, False)
            o = torch.empty(self.grid_shape, dtype=torch.float64, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_ylm_real(self._handle, f.ptr, g.ptr, C.c_void_p(o.data_ptr()),
                                                                  _stream_ptr(self.device)))
            return self._finish(o, f_lm, out)

    def cross_correlation_ylm_real_3d(self, f_lms, g_lms, out=None, radial_sampling_points=None,
                                      radial_limits=None, use_mp=False):
        """Radially weighted correlation of real functions given per shell (pysofft/soft.py:666-703)."""
        torch = _torch()
        n = len(f_lms)
        if radial_sampling_points is None:
            radial_sampling_points = np.linspace(0, 1, n)
        if radial_limits is None:
            radial_limits = [0, n]
        rad = np.ascontiguousarray(radial_sampling_points, dtype=np.float64)
        with torch.cuda.device(self.device):
            f = self._mlr_sets(f_lms, True)
            g = self._mlr_sets(g_lms, True)
            o = torch.empty(self.grid_shape, dtype=torch.float64, device=self.device)
            _lib.check(_lib.lib().so3ft_cross_correlation_ylm_real_3d(
                self._handle, f.ptr, g.ptr, C.c_void_p(o.data_ptr()),
                rad.ctypes.data_as(C.POINTER(C.c_double)), len(rad), int(radial_limits[0]),
                int(radial_limits[1]) - 1, _stream_ptr(self.device)))
            return self._finish(o, f_lms, out)

    def _inverse_wigner_trf_corr_cmplx(self, f_lm, g_lm, out=None, use_mp=False):
        """Correlation DWT stage (pysofft/soft.py:550-554): the fft array before the FFT."""
        torch = _torch()
        nb = self.bw ** 2
        with torch.cuda.device(self.device):
            f = _Dev(f_lm, torch.complex128, self.device, (nb,))
            g = _Dev(g_lm, torch.complex128, self.device, (nb,))
            o = torch.empty(self.grid_shape, dtype=torch.complex128, device=self.device)
            _lib.check(_lib.lib().so3ft_inverse_wigner_trf_corr_cmplx(self._handle, f.ptr, g.ptr,
                                                                     C.c_void_p(o.data_ptr()),
                                                                     _stream_ptr(self.device)))
            return self._finish(o, f_lm, out)

    def _inverse_wigner_trf_corr_real(self, f_ml, g_ml, out=None, use_mp=False):
        """Real correlation DWT stage (pysofft/soft.py:555-559): half-complex fft array."""
        torch = _torch()
        with torch.cuda.device(self.device):
            f = self._mlr_sets(f_ml, False)
            g = self._mlr_sets(g_ml, False)
            o = torch.empty(self.grid_shape, dtype=torch.complex128, device=self.device)
            _lib.check(_lib.lib().so3ft_inverse_wigner_trf_corr_real(self._handle, f.ptr, g.ptr,
                                                                    C.c_void_p(o.data_ptr()),
                                                                    _stream_ptr(self.device)))
            return self._finish(o, f_ml, out)

    def _finish(self, o, like, out):
        if _is_torch(like):
            if out is not None:
                out.copy_(o)
                return out
            return o
        res = o.cpu().numpy()
        if out is None:
            return res
        np.copyto(out, res)
        return out

    # -- rotations of SH coefficients (pysofft/soft.py:713-807) ------------------
    @staticmethod
    def rotate_ylm_cmplx(ylms, euler_angles):
        from .rotate import rotate_ylm_cmplx
        return rotate_ylm_cmplx(ylms, euler_angles)

    @staticmethod
    def rotate_ylm_real(ymls, euler_angles):
        from .rotate import rotate_ylm_real
        return rotate_ylm_real(ymls, euler_angles)


class CoeffSO3(np.ndarray):
    r"""ndarray of Wigner coefficients with ``.lmn[l, m, n]`` access (pysofft/soft.py:811-908)."""
    _coeff_orders = ['mnl', 'lmn']

    def __new__(cls, coeff, lmn_indices, coeff_order='mnl'):
        obj = np.asarray(coeff).view(cls)
        lmn_indices.flags.writeable = False
        obj._lmns = lmn_indices
        obj.lmn = CoeffView(obj, lmn_indices, coeff_order)
        obj.bw = obj.lmn.bw
        obj.coeff_order = obj.lmn.coeff_order
        return obj

    def __array_finalize__(self, obj):
        if obj is None:
            return
        self._lmns = getattr(obj, '_lmns', None)
        if self._lmns is None:
            self.lmn = None
            self.bw = None
            self.coeff_order = None
        else:
            self.lmn = CoeffView(self, self._lmns, getattr(obj, 'coeff_order', 'mnl') or 'mnl')
            self.bw = self.lmn.bw
            self.coeff_order = self.lmn.coeff_order

    def __array_ufunc__(self, ufunc, method, *inputs, out=None, **kwargs):
        args = [i.view(np.ndarray) if isinstance(i, CoeffSO3) else i for i in inputs]
        outputs = out
        if outputs:
            kwargs['out'] = tuple(o.view(np.ndarray) if isinstance(o, CoeffSO3) else o for o in outputs)
        else:
            outputs = (None,) * ufunc.nout
        results = super().__array_ufunc__(ufunc, method, *args, **kwargs)
        if results is NotImplemented:
            return NotImplemented
        if method == 'at':
            return None
        if ufunc.nout == 1:
            results = (results,)
        results = list((np.asarray(r).view(CoeffSO3) if o is None else o) for r, o in zip(results, outputs))
        coeff_inputs = tuple(i for i in inputs if isinstance(i, CoeffSO3))
        for k, res in enumerate(results):
            done = False
            if isinstance(res, CoeffSO3):
                for inp in coeff_inputs:
                    if res.shape == inp.shape:
                        res._lmns = inp._lmns
                        res.lmn = CoeffView(res, res._lmns, inp.coeff_order)
                        res.bw = res.lmn.bw
                        res.coeff_order = inp.coeff_order
                        done = True
                        break
                if not done:
                    results[k] = res.view(np.ndarray)
        return results[0] if len(results) == 1 else results


class CoeffView:
    """``view[l, m, n]`` access to f^l_{m,n} (pysofft/soft.py:910-976)."""
    _coeff_orders = ['mnl', 'lmn']

    def __init__(self, coeff_array, lmn_indices, coeff_order='mnl'):
        self.array = coeff_array
        self._lmns = lmn_indices
        self.ls = np.sort(np.unique(lmn_indices[:, 0]))
        bw = len(self.ls)
        self.bw = bw
        self.ns = np.roll(np.sort(np.unique(lmn_indices[:, 1])), bw)
        self.ks = np.roll(np.sort(np.unique(lmn_indices[:, 2])), bw)
        if coeff_order not in self._coeff_orders:
            raise AssertionError(f'provided coeff_order {coeff_order} not contained in {self._coeff_orders}.')
        self.coeff_order = coeff_order

    def _check_index(self, l, m, n):
        assert 0 <= l < self.bw, f'l={l} does not satisfy 0<=l<bw={self.bw}'
        mn = max(abs(m), abs(n))
        assert l >= mn, f'l={l} does not satisfy  l>=Max(|m|,|n|)={mn}'

    def _get_mask(self, items):
        if not isinstance(items, tuple):
            items = (items,)
        assert len(items) <= 3
        if len(items) == 3 and all(isinstance(i, (int, np.integer)) for i in items):
            self._check_index(items[0], items[1], items[2])
            if self.coeff_order == 'lmn':
                id_ = utils.coeff_location_lmn(items[0], items[1], items[2]) - 1
            else:
                id_ = utils.coeff_location_mnl(items[1], items[2], items[0], self.bw) - 1
            return slice(id_, id_ + 1)
        selection_ids = [ids[sel] for sel, ids in zip(items, [self.ls, self.ns, self.ks])]
        masks = [np.isin(lmn, sids) for sids, lmn in zip(selection_ids, self._lmns.T)]
        return np.prod(masks, axis=0, dtype=bool)

    def __getitem__(self, items):
        return self.array[..., self._get_mask(items)]

    def __setitem__(self, items, value):
        self.array[..., self._get_mask(items)] = value
