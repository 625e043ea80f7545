"""GPU rotation of spherical-harmonic coefficients (pysofft/soft.py:713-807)."""
import ctypes as C

import numpy as np

from . import _lib, utils

_lib.SIGNATURES["so3ft_rotate_ylm_cmplx"] = (C.c_int, [C.c_int64, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                                      C.c_int, C.c_void_p, C.c_void_p])


def _rot_fn():
    L = _lib.lib()
    f = L.so3ft_rotate_ylm_cmplx
    f.restype, f.argtypes = _lib.SIGNATURES["so3ft_rotate_ylm_cmplx"]
    return f


def rotate_ylm_device(ylms, euler, paired=False):
    """torch CUDA in/out: ylms (nset, bw^2) complex128, euler (nrot, 3) float64.
    Returns (nrot, nset, bw^2), or (nrot, bw^2) when ``paired`` (set i by rotation i)."""
    import torch
    y = ylms.to(torch.complex128).contiguous()
    e = euler.to(torch.float64).contiguous()
    nset, nb = y.shape
    nrot = e.shape[0]
    bw = int(round(np.sqrt(nb)))
    if bw * bw != nb:
        raise ValueError("coefficient sets must have bw^2 entries")
    shape = (nrot, nb) if paired else (nrot, nset, nb)
    out = torch.empty(shape, dtype=torch.complex128, device=y.device)
    s = C.c_void_p(torch.cuda.current_stream(y.device).cuda_stream)
    _lib.check(_rot_fn()(bw, C.c_void_p(y.data_ptr()), nset, C.c_void_p(e.data_ptr()), nrot, int(bool(paired)),
                         C.c_void_p(out.data_ptr()), s))
    return out


def rotate_ylm_cmplx(ylms, euler_angles):
    """soft.py:713-764 semantics: (m rotations, n sets) -> (m, n, bw^2), squeezed when m or n is 1."""
    import torch
    is_t = isinstance(ylms, torch.Tensor)
    dev = ylms.device if (is_t and ylms.is_cuda) else torch.device("cuda", torch.cuda.current_device())
    y = ylms if is_t else torch.from_numpy(np.atleast_2d(np.asarray(ylms, dtype=np.complex128)).copy())
    if y.ndim == 1:
        y = y[None]
    e = euler_angles if isinstance(euler_angles, torch.Tensor) else torch.from_numpy(
        np.atleast_2d(np.asarray(euler_angles, dtype=np.float64)).copy())
    if e.ndim == 1:
        e = e[None]
    out = rotate_ylm_device(y.to(dev), e.to(dev))
    if e.shape[0] == 1 or y.shape[0] == 1:
        out = out.squeeze()
    return out if is_t else out.cpu().numpy()


def rotate_ylm_real(ymls, euler_angles):
    """soft.py:766-807: MLr-ordered real-function coefficients via the complex rotation."""
    ymls = np.atleast_2d(np.asarray(ymls))
    bw = int((np.sqrt(1 + 8 * ymls.shape[-1]) - 1) / 2)
    ylms = np.zeros(ymls.shape[:-1] + (bw ** 2,), dtype=complex)
    for l in range(bw):
        for m in range(l + 1):
            ml_id = utils.mlr(m, l, bw) - 1
            lm_id = utils.lmc(l, m) - 1
            lm_neg_id = utils.lmc(l, -m) - 1
            ylms[..., lm_id] = ymls[..., ml_id]
            ylms[..., lm_neg_id] = ylms[..., lm_id].conj() * (-1) ** m
    ylms_rot = rotate_ylm_cmplx(ylms, euler_angles)
    ymls_rot = np.zeros(ylms_rot.shape[:-1] + (ymls.shape[-1],), dtype=complex)
    for m in range(bw):
        for l in range(m, bw):
            ymls_rot[..., utils.mlr(m, l, bw) - 1] = ylms_rot[..., utils.lmc(l, m) - 1]
    return ymls_rot
