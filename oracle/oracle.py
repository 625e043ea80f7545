"""ctypes bindings for the C oracle (liboracle.so) and the compiled reference
modules (_ref/libsoftref.so).  TEST INFRASTRUCTURE ONLY (see package docstring).

Array conventions match the reference's Python API (pysofft/soft.py):
coefficients are 1-D complex128 in mnl order; SO(3) functions are C-order
(gamma, alpha, beta) arrays of shape (2bw, 2bw, 2bw).
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_REF = os.path.join(_HERE, "_ref", "libsoftref.so")

_i64 = C.c_int64
_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)
_vp = C.c_void_p


def build():
    """Compile liboracle.so (and _ref when the reference checkout exists)."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def _load():
    if not os.path.exists(_LIB):
        build()
    lib = C.CDLL(_LIB)
    sig = {
        "or_total_num_coeffs": (_i64, [_i64]),
        "or_slice_table": (None, [_i64, _ip]),
        "or_coeff_degrees": (None, [_i64, _ip]),
        "or_beta_samples": (None, [_i64, _dp]),
        "or_legendre_weights": (None, [_i64, _dp]),
        "or_seeds": (None, [_i64, _i64, _dp, _dp, _dp]),
        "or_genwig": (None, [_i64, _i64, _i64, _i64, _dp, _dp, _dp]),
        "or_plan_create": (_vp, [_i64, _i64]),
        "or_plan_destroy": (None, [_vp]),
        "or_plan_weights": (None, [_vp, _dp]),
        "or_plan_seeds": (None, [_vp, _dp]),
        "or_inverse_dwt": (None, [_vp, _vp, _vp, C.c_int]),
        "or_forward_dwt": (None, [_vp, _vp, _vp, C.c_int]),
        "or_inverse_dwt_real": (None, [_vp, _vp, _vp, C.c_int]),
        "or_forward_dwt_real": (None, [_vp, _vp, _vp, C.c_int]),
        "or_fft2": (None, [_vp, _vp, C.c_int, C.c_int]),
        "or_isoft": (None, [_vp, _vp, _vp, C.c_int]),
        "or_soft": (None, [_vp, _vp, _vp, C.c_int]),
        "or_irsoft": (None, [_vp, _vp, _vp, C.c_int]),
        "or_rsoft": (None, [_vp, _vp, _vp, C.c_int]),
        "or_isoft_many": (None, [_vp, _vp, _vp, _i64, C.c_int]),
        "or_soft_many": (None, [_vp, _vp, _vp, _i64, C.c_int]),
        "or_irsoft_many": (None, [_vp, _vp, _vp, _i64, C.c_int]),
        "or_rsoft_many": (None, [_vp, _vp, _vp, _i64, C.c_int]),
        "or_integrate": (None, [_vp, _vp, _dp, _dp]),
        "or_enforce_real_sym": (None, [_i64, _vp]),
        "or_corr_coeffs": (None, [_i64, _vp, _vp, _vp]),
        "or_risbo_reduced_step": (None, [_dp, _i64, _i64, _dp, _dp, _dp, C.c_int, _dp]),
        "or_dtab_rows": (_i64, [_i64]),
        "or_genwig_all_risbo": (None, [_i64, C.c_int, _dp]),
        "or_genwig_all_kostelec": (None, [_i64, _dp]),
        "or_full_big_d_risbo": (None, [_i64, C.c_double, C.c_double, C.c_double, C.c_int, _vp]),
        "or_cross_correlation_cmplx": (None, [_vp, _vp, _vp, _vp, C.c_int]),
        "or_cross_correlation_cmplx_3d": (None, [_vp, _vp, _vp, _i64, _dp, _i64, _i64, _vp, C.c_int]),
        "or_risbo_step": (None, [_dp, _i64, C.c_double, C.c_double, _dp, _dp]),
        "or_rotate_ylm_cmplx": (None, [_i64, _vp, _i64, _dp, _i64, _vp, C.c_int]),
        "or_num_threads": (C.c_int, []),
        "or_set_num_threads": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_lib = _load()


def _p(a, t=_dp):
    return a.ctypes.data_as(t)


def _v(a):
    return C.c_void_p(a.ctypes.data)


def _c128(a):
    return np.ascontiguousarray(a, dtype=np.complex128)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# index / sampling helpers
# --------------------------------------------------------------------------
def total_num_coeffs(bw):
    return int(_lib.or_total_num_coeffs(bw))


def slice_table(bw):
    """0-based start of f^l_{m,n} in the mnl layout, indexed [m+bw-1, n+bw-1]."""
    t = np.zeros((2 * bw - 1) * (2 * bw - 1), np.int64)
    _lib.or_slice_table(bw, _p(t, _ip))
    return t.reshape(2 * bw - 1, 2 * bw - 1)


def coeff_degrees(bw):
    out = np.zeros((total_num_coeffs(bw), 3), np.int64)
    _lib.or_coeff_degrees(bw, _p(out, _ip))
    return out


def beta_samples(n):
    b = np.zeros(n)
    _lib.or_beta_samples(n, _p(b))
    return b


def legendre_weights(bw):
    w = np.zeros(2 * bw)
    _lib.or_legendre_weights(bw, _p(w))
    return w


def seeds(bw):
    N = 2 * bw
    beta = beta_samples(N)
    sc = np.cos(beta / 2) * np.sin(beta / 2)
    c2 = np.cos(beta / 2) ** 2
    out = np.zeros((bw * (bw + 1) // 2, N))
    _lib.or_seeds(bw, N, _p(sc), _p(c2), _p(out))
    return out


def genwig(p, q, bw):
    """d~^l_{p,q}(beta_k), l = q..bw-1 on the bw grid, shape (bw-q, 2bw)."""
    N = 2 * bw
    s = seeds(bw)
    tri = (p * (2 * bw - p + 1)) // 2 + q - p
    seed = np.ascontiguousarray(s[tri])
    cosb = np.cos(beta_samples(N))
    out = np.zeros((bw - q, N))
    _lib.or_genwig(p, q, bw, N, _p(cosb), _p(seed), _p(out))
    return out


def mlr_to_lmc(fml, bw):
    """MLr coefficients of a real function -> LMc (index l(l+1)+m) with
    f^l_{-m} = (-1)^m conj(f^l_m)."""
    fml = np.asarray(fml, np.complex128)
    out = np.zeros(fml.shape[:-1] + (bw * bw,), np.complex128)
    for l in range(bw):
        for m in range(l + 1):
            v = fml[..., (m * (2 * bw - 1 - m)) // 2 + l]
            out[..., l * (l + 1) + m] = v
            out[..., l * (l + 1) - m] = (-1) ** m * np.conj(v)
    return out


def enforce_real_sym(coeff, bw):
    """In-place, utils.f90:471-488."""
    assert coeff.dtype == np.complex128 and coeff.flags.c_contiguous
    _lib.or_enforce_real_sym(bw, _v(coeff))
    return coeff


def corr_coeffs(bw, flm, glm):
    cc = np.zeros(total_num_coeffs(bw), np.complex128)
    _lib.or_corr_coeffs(bw, _v(_c128(flm)), _v(_c128(glm)), _v(cc))
    return cc


def set_num_threads(n):
    _lib.or_set_num_threads(int(n))


def num_threads():
    return int(_lib.or_num_threads())


def risbo_step(dprev, l, c, s, sq):
    out = np.zeros((2 * l + 1) ** 2)
    dp = _f64(dprev) if dprev is not None and len(dprev) else np.zeros(1)
    _lib.or_risbo_step(_p(dp), l, c, s, _p(_f64(sq)), _p(out))
    return out


def risbo_reduced_step(d1, l, c, s, normalized=True, sq=None):
    """wigner.f90:626-748: one step d^{l-1} -> d^l of the reduced Risbo recurrence at the
    betas whose half-angle cosines / sines are c, s (a set closed under beta -> pi - beta,
    reversed); d1, result: [tri][nb]."""
    c, s = _f64(c), _f64(s)
    nb = len(c)
    sq = _f64(np.sqrt(np.arange(max(3, 2 * l + 2))) if sq is None else sq)
    out = np.zeros(((l + 1) * (l + 2) // 2, nb))
    dp = _f64(d1) if d1 is not None and np.size(d1) else np.zeros(nb)
    _lib.or_risbo_reduced_step(_p(dp), l, nb, _p(c), _p(s), _p(sq), int(normalized), _p(out))
    return out


def dtab_rows(bw):
    """rows (p, q, l), 0 <= p <= q <= l < bw, of the mnl-ordered d table"""
    return int(_lib.or_dtab_rows(bw))


def genwig_all_risbo(bw, normalized=True):
    """wigner.f90:784-822: the d table by the Risbo recurrence, [mnl row][2bw]"""
    out = np.zeros((dtab_rows(bw), 2 * bw))
    _lib.or_genwig_all_risbo(bw, int(normalized), _p(out))
    return out


def genwig_all_kostelec(bw):
    """wigner.f90:423-465: the d table by the Kostelec recurrence, [mnl row][2bw]"""
    out = np.zeros((dtab_rows(bw), 2 * bw))
    _lib.or_genwig_all_kostelec(bw, _p(out))
    return out


def full_big_d_risbo(bw, alpha, beta, gamma, normalized=True):
    """wigner.f90:844-934: D^l_{mn}(alpha, beta, gamma) in mnl order (complex128)"""
    out = np.zeros(total_num_coeffs(bw), np.complex128)
    _lib.or_full_big_d_risbo(bw, alpha, beta, gamma, int(normalized), _v(out))
    return out


def rotate_ylm_cmplx(ylms, euler_angles, use_mp=True):
    """soft.py:713-764 semantics; returns (nrot, nset, bw^2) (no squeeze)."""
    y = np.atleast_2d(_c128(ylms))
    e = np.atleast_2d(_f64(euler_angles))
    bw = int(round(np.sqrt(y.shape[-1])))
    out = np.zeros((e.shape[0], y.shape[0], bw * bw), np.complex128)
    _lib.or_rotate_ylm_cmplx(bw, _v(y), y.shape[0], _p(e), e.shape[0], _v(out), int(use_mp))
    return out


class OraclePlan:
    """FP64 CPU restatement of one so3ft instance (soft.f08:16-126)."""

    def __init__(self, bw, lmax=None):
        self.bw = int(bw)
        self.lmax = self.bw - 1 if lmax is None else int(lmax)
        self.N = 2 * self.bw
        self.nc = total_num_coeffs(self.bw)
        self._h = _lib.or_plan_create(self.bw, self.lmax)

    def __del__(self):
        if getattr(self, "_h", None):
            _lib.or_plan_destroy(self._h)
            self._h = None

    @property
    def grid_shape(self):
        return (self.N, self.N, self.N)

    def weights(self):
        w = np.zeros(self.N)
        _lib.or_plan_weights(self._h, _p(w))
        return w

    def isoft(self, coeff, use_mp=True):
        c = _c128(coeff)
        out = np.zeros(self.grid_shape, np.complex128)
        _lib.or_isoft(self._h, _v(c), _v(out), int(use_mp))
        return out

    def soft(self, f, use_mp=True):
        f = _c128(f)
        out = np.zeros(self.nc, np.complex128)
        _lib.or_soft(self._h, _v(f), _v(out), int(use_mp))
        return out

    def irsoft(self, coeff, use_mp=True):
        c = _c128(coeff)
        out = np.zeros(self.grid_shape, np.float64)
        _lib.or_irsoft(self._h, _v(c), _v(out), int(use_mp))
        return out

    def rsoft(self, f, use_mp=True):
        f = _f64(f)
        out = np.zeros(self.nc, np.complex128)
        _lib.or_rsoft(self._h, _v(f), _v(out), int(use_mp))
        return out

    def isoft_many(self, coeffs, use_mp=True):
        c = _c128(coeffs)
        out = np.zeros((c.shape[0],) + self.grid_shape, np.complex128)
        _lib.or_isoft_many(self._h, _v(c), _v(out), c.shape[0], int(use_mp))
        return out

    def soft_many(self, fs, use_mp=True):
        f = _c128(fs)
        out = np.zeros((f.shape[0], self.nc), np.complex128)
        _lib.or_soft_many(self._h, _v(f), _v(out), f.shape[0], int(use_mp))
        return out

    def irsoft_many(self, coeffs, use_mp=True):
        c = _c128(coeffs)
        out = np.zeros((c.shape[0],) + self.grid_shape, np.float64)
        _lib.or_irsoft_many(self._h, _v(c), _v(out), c.shape[0], int(use_mp))
        return out

    def rsoft_many(self, fs, use_mp=True):
        f = _f64(fs)
        out = np.zeros((f.shape[0], self.nc), np.complex128)
        _lib.or_rsoft_many(self._h, _v(f), _v(out), f.shape[0], int(use_mp))
        return out

    def inverse_dwt(self, coeff, use_mp=True):
        """Stage: DWT output array (i_n, i_m, beta) before the FFT (zeros elsewhere)."""
        out = np.zeros(self.grid_shape, np.complex128)
        _lib.or_inverse_dwt(self._h, _v(_c128(coeff)), _v(out), int(use_mp))
        return out

    def forward_dwt(self, F, use_mp=True):
        out = np.zeros(self.nc, np.complex128)
        _lib.or_forward_dwt(self._h, _v(_c128(F)), _v(out), int(use_mp))
        return out

    def fft2(self, X, sign, use_mp=True):
        X = _c128(X).copy()
        _lib.or_fft2(self._h, _v(X), int(sign), int(use_mp))
        return X

    def integrate(self, f):
        f = _c128(f)
        re, im = C.c_double(), C.c_double()
        _lib.or_integrate(self._h, _v(f), C.byref(re), C.byref(im))
        return complex(re.value, im.value)

    def cross_correlation_ylm_cmplx(self, flm, glm, use_mp=True):
        out = np.zeros(self.grid_shape, np.complex128)
        _lib.or_cross_correlation_cmplx(self._h, _v(_c128(flm)), _v(_c128(glm)), _v(out), int(use_mp))
        return out

    def cross_correlation_ylm_real(self, fml, gml, stage=False):
        """Real correlation from MLr sets (index m(2bw-1-m)/2 + l), restated loop by loop
        from soft.f08:2806-2872 (half-complex DWT of the pairs 0 <= m1 <= m2), :2932-2945
        (the (0, -n) columns by conjugation) and :3118-3131 (conj, c2r FFTW_BACKWARD,
        /2pi, here written as the real part of the forward 2-D DFT of the Hermitian
        completion).  stage=True returns the half-complex fft array (i_n, i_m, beta)
        before the FFT (py_inverse_wigner_trf_corr_real).  Pure numpy over pairs: small
        bandwidths only."""
        bw, N = self.bw, self.N
        fml = np.asarray(fml, np.complex128)
        gml = np.asarray(gml, np.complex128)
        lmax = self.lmax
        F = np.zeros((N, N, N), np.complex128)          # (i_n, i_m, beta)
        l_all = np.arange(bw)
        wn_all = 2 * np.pi * np.sqrt(2.0 / (2 * l_all + 1))
        sym_all = (-1.0) ** l_all
        mlr = lambda m, ls: (m * (2 * bw - 1 - m)) // 2 + ls
        for m1 in range(lmax + 1):
            s1 = (-1.0) ** m1
            for m2 in range(m1, lmax + 1):
                s2 = (-1.0) ** m2
                ls = np.arange(m2, bw)
                wig = genwig(m1, m2, bw)                    # (bw - m2, N)
                wn, sy = wn_all[m2:], sym_all[m2:]
                f1, f2 = fml[mlr(m1, ls)], fml[mlr(m2, ls)]
                g1, g2 = gml[mlr(m1, ls)], gml[mlr(m2, ls)]
                F[m2, m1, :] = (wn * f1 * np.conj(g2) * s1 * s2) @ wig
                if m1 == 0 and m2 == 0:
                    continue
                if m1 != m2:
                    F[m1, m2, :] = (wn * f2 * np.conj(g1)) @ wig
                if m1 == 0 or m2 == 0:
                    continue
                F[N - m2, m1, ::-1] = s1 * ((wn * f1 * g2 * s1 * sy) @ wig)
                if m1 == m2:
                    continue
                F[N - m1, m2, ::-1] = s1 * ((wn * f2 * g1 * s2 * sy) @ wig)
        for m2 in range(1, lmax + 1):
            F[N - m2, 0, :] = np.conj(F[m2, 0, :])
        if stage:
            return F
        # Hermitian completion of the m < 0 columns: X(-m, -n) = conj X(m, n)
        for im in range(bw + 1, N):
            for i_n in range(N):
                F[i_n, im, :] = np.conj(F[(N - i_n) % N, N - im, :])
        return np.fft.fft2(F, axes=(0, 1)).real / (2 * np.pi)

    def cross_correlation_ylm_real_3d(self, fmls, gmls, radial_sampling_points, radial_limits):
        """soft.f08:3164-3253: the radially weighted sum of the shells' real correlations,
        radial_limits = (lo, hi) with Python semantics [lo, hi)."""
        r = np.asarray(radial_sampling_points, np.float64)
        lo, hi = int(radial_limits[0]), int(radial_limits[1]) - 1
        inv = 1.0 / (r[hi] - r[lo] + (r[1] - r[0]))
        out = np.zeros(self.grid_shape)
        for k in range(lo, hi + 1):
            out += r[k] ** 2 * inv * self.cross_correlation_ylm_real(fmls[k], gmls[k])
        return out

    def cross_correlation_ylm_cmplx_3d(self, flms, glms, radial_sampling_points, radial_limits, use_mp=True):
        """radial_limits = (lo, hi) with Python semantics [lo, hi) as soft.py:660-664."""
        f = _c128(flms)
        g = _c128(glms)
        r = _f64(radial_sampling_points)
        out = np.zeros(self.grid_shape, np.complex128)
        lo, hi = int(radial_limits[0]), int(radial_limits[1]) - 1
        _lib.or_cross_correlation_cmplx_3d(self._h, _v(f), _v(g), f.shape[0], _p(r), lo, hi, _v(out), int(use_mp))
        return out


# --------------------------------------------------------------------------
# the reference's own utils/make_wigner modules (compiled in place)
# --------------------------------------------------------------------------
class RefLib:
    """Thin wrapper over _ref/libsoftref.so; ``RefLib.available()`` tells whether it was built."""

    _h = None

    @classmethod
    def available(cls):
        return os.path.exists(_REF)

    @classmethod
    def lib(cls):
        if cls._h is None:
            h = C.CDLL(_REF)
            sig = {
                "ref_total_num_coeffs": (_i64, [_i64]),
                "ref_coeff_location_mnl": (_i64, [_i64, _i64, _i64, _i64]),
                "ref_coeff_degrees": (None, [_i64, _ip]),
                "ref_legendre_weights": (None, [_i64, _dp]),
                "ref_beta_samples": (None, [_i64, _dp]),
                "ref_alpha_gamma_samples": (None, [_i64, _dp]),
                "ref_dlml_all": (None, [_i64, _dp]),
                "ref_genwig": (None, [_i64, _i64, _i64, _dp]),
                "ref_wigner_dl_kostelec": (None, [_i64, _i64, _dp, C.c_int, _dp]),
                "ref_risbo_step": (None, [_dp, _i64, C.c_double, C.c_double, _dp, _i64, _dp]),
                "ref_full_big_d": (None, [_i64, C.c_double, C.c_double, C.c_double, _vp]),
                "ref_full_big_d_norm": (None, [_i64, C.c_double, C.c_double, C.c_double, C.c_int, _vp]),
                "ref_risbo_reduced_step": (None, [_dp, _i64, _i64, _dp, _dp, _dp, _i64, C.c_int, _dp]),
                "ref_genwig_all_risbo": (None, [_i64, C.c_int, _dp]),
                "ref_genwig_all_kostelec": (None, [_i64, _dp]),
                "ref_enforce_real_sym_mnl": (None, [_vp, _i64]),
                "ref_coeff_location_lmn": (_i64, [_i64, _i64, _i64]),
                "ref_coeff_degrees_risbo": (None, [_i64, _ip]),
                "ref_enforce_real_sym_lmn": (None, [_vp, _i64]),
                "ref_lmc": (_i64, [_i64, _i64]),
                "ref_mlr": (_i64, [_i64, _i64, _i64]),
                "ref_mlc": (_i64, [_i64, _i64, _i64]),
            }
            for name, (res, args) in sig.items():
                f = getattr(h, name)
                f.restype = res
                f.argtypes = args
            cls._h = h
        return cls._h

    @classmethod
    def coeff_location_mnl(cls, m, n, l, bw):
        return int(cls.lib().ref_coeff_location_mnl(m, n, l, bw)) - 1

    @classmethod
    def coeff_degrees(cls, bw):
        out = np.zeros((total_num_coeffs(bw), 3), np.int64)
        cls.lib().ref_coeff_degrees(bw, _p(out, _ip))
        return out

    @classmethod
    def coeff_location_lmn(cls, l, m, n):
        return int(cls.lib().ref_coeff_location_lmn(l, m, n)) - 1

    @classmethod
    def coeff_degrees_risbo(cls, bw):
        out = np.zeros((total_num_coeffs(bw), 3), np.int64)
        cls.lib().ref_coeff_degrees_risbo(bw, _p(out, _ip))
        return out

    @classmethod
    def enforce_real_sym_lmn(cls, coeff, bw):
        c = np.ascontiguousarray(coeff, np.complex128).copy()
        cls.lib().ref_enforce_real_sym_lmn(_v(c), bw)
        return c

    @classmethod
    def legendre_weights(cls, bw):
        w = np.zeros(2 * bw)
        cls.lib().ref_legendre_weights(bw, _p(w))
        return w

    @classmethod
    def beta_samples(cls, n):
        b = np.zeros(n)
        cls.lib().ref_beta_samples(n, _p(b))
        return b

    @classmethod
    def alpha_gamma_samples(cls, n):
        b = np.zeros(n)
        cls.lib().ref_alpha_gamma_samples(n, _p(b))
        return b

    @classmethod
    def seeds(cls, bw):
        out = np.zeros((bw * (bw + 1) // 2, 2 * bw))
        cls.lib().ref_dlml_all(bw, _p(out))
        return out

    @classmethod
    def genwig(cls, p, q, bw):
        out = np.zeros((bw - q, 2 * bw))
        cls.lib().ref_genwig(p, q, bw, _p(out))
        return out

    @classmethod
    def wigner_dl_kostelec(cls, l, betas, normalized=False):
        b = _f64(betas)
        out = np.zeros((len(b), 2 * l + 1, 2 * l + 1))
        cls.lib().ref_wigner_dl_kostelec(l, len(b), _p(b), int(normalized), _p(out))
        return out  # [beta][n+l][m+l] (Fortran dl(beta, n, m))

    @classmethod
    def risbo_step(cls, dprev, l, c, s, sq):
        out = np.zeros((2 * l + 1) ** 2)
        dp = _f64(dprev) if dprev is not None and len(dprev) else np.zeros(1)
        sq = _f64(sq)
        cls.lib().ref_risbo_step(_p(dp), l, c, s, _p(sq), len(sq), _p(out))
        return out

    @classmethod
    def full_big_d(cls, bw, alpha, beta, gamma):
        out = np.zeros(total_num_coeffs(bw), np.complex128)
        cls.lib().ref_full_big_d(bw, alpha, beta, gamma, _v(out))
        return out

    @classmethod
    def full_big_d_norm(cls, bw, alpha, beta, gamma, normalized=True):
        out = np.zeros(total_num_coeffs(bw), np.complex128)
        cls.lib().ref_full_big_d_norm(bw, alpha, beta, gamma, int(normalized), _v(out))
        return out

    @classmethod
    def risbo_reduced_step(cls, d1, l, c, s, normalized=True, sq=None):
        c, s = _f64(c), _f64(s)
        nb = len(c)
        sq = _f64(np.sqrt(np.arange(max(3, 2 * l + 2))) if sq is None else sq)
        out = np.zeros(((l + 1) * (l + 2) // 2, nb))
        dp = _f64(d1) if d1 is not None and np.size(d1) else np.zeros(nb)
        cls.lib().ref_risbo_reduced_step(_p(dp), l, nb, _p(c), _p(s), _p(sq), len(sq), int(normalized), _p(out))
        return out

    @classmethod
    def genwig_all_risbo(cls, bw, normalized=True):
        out = np.zeros((dtab_rows(bw), 2 * bw))
        cls.lib().ref_genwig_all_risbo(bw, int(normalized), _p(out))
        return out

    @classmethod
    def genwig_all_kostelec(cls, bw):
        out = np.zeros((dtab_rows(bw), 2 * bw))
        cls.lib().ref_genwig_all_kostelec(bw, _p(out))
        return out

    @classmethod
    def enforce_real_sym_mnl(cls, coeff, bw):
        assert coeff.dtype == np.complex128 and coeff.flags.c_contiguous
        cls.lib().ref_enforce_real_sym_mnl(_v(coeff), bw)
        return coeff

    @classmethod
    def lmc(cls, l, m):
        return int(cls.lib().ref_lmc(l, m)) - 1

    @classmethod
    def mlr(cls, m, l, bw):
        return int(cls.lib().ref_mlr(m, l, bw)) - 1

    @classmethod
    def mlc(cls, m, l, bw):
        return int(cls.lib().ref_mlc(m, l, bw)) - 1
