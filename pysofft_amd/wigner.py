"""Helpers of the reference's ``pysofft.wigner`` namespace (the make_wigner module of
pysofft/wigner.f90, exported at pysofft/__init__.py:12-13).

The sampling helpers are closed forms evaluated on the host.  The Wigner-d recurrences
run on the GPU through libso3ft (include/so3ft.h); numpy inputs are staged to the device
and the results returned as numpy arrays, torch CUDA tensors stay on the device.  Array
shapes are the ones the reference returns to Python.
"""
import ctypes as C

import numpy as np

from . import _lib


def create_beta_samples(n):
    """pysofft/wigner.f90:34-39: (2i+1) pi / (2n), i < n."""
    n = int(n)
    return (2 * np.arange(n) + 1) * (np.pi / (2 * n))


def create_alpha_gamma_samples(n):
    """pysofft/wigner.f90:94-100: 2 pi i / n, i < n."""
    n = int(n)
    return np.arange(n) * (2.0 * np.pi / n)


def create_trig_samples(bw):
    """pysofft/wigner.f90:51-58: columns cos(beta), cos(beta/2) sin(beta/2), cos(beta/2)^2."""
    b = create_beta_samples(2 * int(bw))
    return np.stack([np.cos(b), np.cos(b / 2) * np.sin(b / 2), np.cos(b / 2) ** 2], axis=1)


def create_trig_samples_risbo(bw):
    """pysofft/wigner.f90:69-75: columns cos(beta/2), sin(beta/2) of the 2bw beta samples."""
    b = create_beta_samples(2 * int(bw))
    return np.stack([np.cos(b / 2), np.sin(b / 2)], axis=1)


def create_sqrts_risbo(bw):
    """pysofft/wigner.f90:81-88: sqrt(i) for 0 <= i <= 2bw - 2."""
    return np.sqrt(np.arange(2 * int(bw) - 1, dtype=np.float64))


def _torch():
    import torch
    return torch


def _dev(x):
    """(tensor on the current CUDA device, was_torch) for a float64 input."""
    torch = _torch()
    if isinstance(x, torch.Tensor):
        t = x if x.is_cuda else x.cuda()
        return t.to(torch.float64).contiguous(), True
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, dtype=np.float64))).cuda(), False


def _stream():
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def wigner_mn_recurrence_risbo(d_l_1, l, cos_beta, sin_beta, sqrts):
    """pysofft/wigner.f90:529-585: Risbo's recurrence for the full small-d matrix, one degree
    l-1 -> l at every beta: d_l_1 (nbeta, (2l-1)^2) -> (nbeta, (2l+1)^2), each row the matrix
    [i][j] row-major.  cos_beta / sin_beta are the half-angle values the caller chose (the
    reference's rotation passes cos(beta/2), sin(-beta/2)); sqrts[k] = sqrt(k)."""
    torch = _torch()
    l = int(l)
    c, was_t = _dev(cos_beta)
    s, _ = _dev(sin_beta)
    sq, _ = _dev(sqrts)
    nb = c.shape[0]
    d = _dev(d_l_1)[0] if l >= 2 else None
    if d is not None and d.numel() < nb * (2 * l - 1) ** 2:
        raise ValueError(f"d_l_1 must hold ({nb}, {(2 * l - 1) ** 2}) values")
    out = torch.empty((nb, (2 * l + 1) ** 2), dtype=torch.float64, device=c.device)
    _lib.check(_lib.lib().so3ft_wigner_mn_recurrence_risbo(_ptr(d), l, _ptr(c), _ptr(s), _ptr(sq), sq.numel(), nb,
                                                           _ptr(out), _stream()))
    return out if was_t else out.cpu().numpy()


def wigner_recurrence_risbo_reduced(dl_1, l, cos_beta, sin_beta, sqrts, normalized=True):
    """pysofft/wigner.f90:626-748: one step of the symmetry-reduced Risbo recurrence,
    d^{l-1}_{m,n} (0 <= m <= n <= l-1) -> d^l_{m,n}: dl_1 (nbeta, l(l+1)/2) ->
    (nbeta, (l+1)(l+2)/2), columns in triangular order (m outer, n inner).  The betas must
    be closed under beta -> pi - beta in reversed order (the step reads pi - beta as the
    reversed array), as the reference's samples are."""
    torch = _torch()
    l = int(l)
    c, was_t = _dev(cos_beta)
    s, _ = _dev(sin_beta)
    sq, _ = _dev(sqrts)
    nb = c.shape[0]
    d = None
    if l >= 2:
        d = _dev(dl_1)[0].reshape(nb, l * (l + 1) // 2).t().contiguous()   # [tri][beta] on the device
    out = torch.empty(((l + 1) * (l + 2) // 2, nb), dtype=torch.float64, device=c.device)
    _lib.check(_lib.lib().so3ft_wigner_recurrence_risbo_reduced(_ptr(d), l, _ptr(c), _ptr(s), _ptr(sq), sq.numel(),
                                                                nb, int(bool(normalized)), _ptr(out), _stream()))
    out = out.t()
    return out if was_t else out.cpu().numpy()


def wigner_dl_risbo_reduced(l, betas, normalized=True):
    """pysofft/wigner.f90:755-775: d^l_{m,n}(beta), 0 <= m <= n <= l, by l+1 steps of the
    reduced recurrence: (nbeta, (l+1)(l+2)/2)."""
    b = np.asarray(betas, dtype=np.float64)
    sq = np.sqrt(np.arange(2 * int(l) + 3, dtype=np.float64))
    d = None
    for i in range(int(l) + 1):
        d = wigner_recurrence_risbo_reduced(d, i, np.cos(b / 2), np.sin(b / 2), sq, normalized)
    return d


def _genwig_all(bw, rt):
    torch = _torch()
    bw = int(bw)
    rows = bw * (2 + 3 * bw + bw * bw) // 6
    out = torch.empty((rows, 2 * bw), dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().so3ft_genwig_all(bw, rt, _ptr(out), _stream()))
    return out


def _degrees_of_rows(bw):
    return np.concatenate([np.arange(q, bw) for p in range(bw) for q in range(p, bw)])


def genwig_all_risbo(bw, normalized=True):
    """pysofft/wigner.f90:812-822: every d~^l_{m,n}(beta_k), 0 <= m <= n <= l < bw, at the 2bw
    beta samples by the reduced Risbo recurrence: (2bw, rows) with rows in mnl order
    (utils.f90:683-695).  normalized=False returns d = d~ / sqrt((2l+1)/2)."""
    out = _genwig_all(bw, 1).cpu().numpy()
    if not normalized:
        out /= np.sqrt((2 * _degrees_of_rows(int(bw)) + 1) / 2.0)[:, None]
    return out.T


def genwig_all_kostelec(bw, normalized=True):
    """pysofft/wigner.f90:473-491: the same table by the Kostelec recurrence, (2bw, rows)."""
    out = _genwig_all(bw, 0).cpu().numpy()
    if not normalized:
        out /= np.sqrt((2 * _degrees_of_rows(int(bw)) + 1) / 2.0)[:, None]
    return out.T


def full_big_d_risbo(bw, alpha, beta, gamma, normalized=True):
    """pysofft/wigner.f90:844-934: D^l_{m,n}(alpha, beta, gamma) = e^{-i m alpha} d^l_{m,n}(beta)
    e^{-i n gamma} for every (l, m, n) in mnl order (complex, (4bw^3-bw)/3), with that
    routine's signs (mixed-sign orders with m + n odd carry an extra -1); normalized: the
    normalised d~ and an overall 1/(2 pi)."""
    torch = _torch()
    bw = int(bw)
    out = torch.empty((4 * bw ** 3 - bw) // 3, dtype=torch.complex128, device="cuda")
    _lib.check(_lib.lib().so3ft_full_big_d_risbo(bw, float(alpha), float(beta), float(gamma), int(bool(normalized)),
                                                 _ptr(out), _stream()))
    return out.cpu().numpy()
