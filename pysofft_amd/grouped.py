"""Mixed-precision SO(3) transforms of a batch of uneven bandwidths.

``SoftGroup(bws)`` fixes one bandwidth per transform.  ``isoft`` / ``soft`` run the
whole batch with complex64 storage (FP64 recurrence and FFT arithmetic, MFMA
contraction with d split into FP32 hi + bfloat16 lo parts) as grouped DWT launches over work lists and
per-bandwidth FFT passes on parallel branches of a HIP graph captured on first use; the
launch shapes and the schedule (one grouped DWT launch, one branch per bandwidth so that
FFT passes overlap other bandwidths' DWTs, or the bandwidths in sequence) are chosen by timing (so3ft_group_* in
include/so3ft.h).  The reference has no counterpart: its ``*_many`` calls
(pysofft/soft.py:255-330) take one bandwidth per ``Soft`` instance.

Arrays are packed in the library's layout: transforms of one bandwidth contiguous,
bandwidths in decreasing order; ``coeff_views`` / ``grid_views`` give per-transform
views in the caller's order (coefficients ``(NC(bw),)`` mnl, functions
``(2bw, 2bw, 2bw)`` C-order (gamma, alpha, beta)).
"""
import ctypes as C

import numpy as np

from . import _lib


def _torch():
    import torch
    return torch


def group_layout(bws):
    """Element offsets (coefficients, grids) of each transform and the packed totals;
    host-only (no GPU needed)."""
    bws = np.ascontiguousarray(bws, dtype=np.int64)
    n = len(bws)
    co = np.zeros(n, np.int64)
    go = np.zeros(n, np.int64)
    ct, gt = C.c_int64(), C.c_int64()
    _lib.check(_lib.lib().so3ft_group_layout(bws.ctypes.data, n, co.ctypes.data, go.ctypes.data,
                                             C.byref(ct), C.byref(gt)))
    return co, go, ct.value, gt.value


class SoftGroup:
    """Grouped complex64 ``isoft`` / ``soft`` over transforms of bandwidths ``bws``."""

    _handle = None

    def __init__(self, bws, device=None):
        torch = _torch()
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.bws = [int(b) for b in bws]
        self.coeff_off, self.grid_off, self.coeff_total, self.grid_total = group_layout(self.bws)
        b = np.ascontiguousarray(self.bws, dtype=np.int64)
        h = C.c_int64()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().so3ft_group_create(b.ctypes.data, len(b), C.byref(h)))
        self._handle = h.value

    def __del__(self):
        if self._handle is not None:
            try:
                _lib.lib().so3ft_group_destroy(self._handle)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._handle = None

    def __len__(self):
        return len(self.bws)

    # ---- packed arrays
    def empty_coeffs(self):
        return _torch().empty(self.coeff_total, dtype=_torch().complex64, device=self.device)

    def empty_grids(self):
        return _torch().empty(self.grid_total, dtype=_torch().complex64, device=self.device)

    def coeff_views(self, flat):
        return [flat[o:o + (4 * b ** 3 - b) // 3] for o, b in zip(self.coeff_off, self.bws)]

    def grid_views(self, flat):
        return [flat[o:o + 8 * b ** 3].view(2 * b, 2 * b, 2 * b) for o, b in zip(self.grid_off, self.bws)]

    def pack_coeffs(self, coeffs):
        """Packed complex64 coefficients from one array per transform (caller's order)."""
        torch = _torch()
        out = self.empty_coeffs()
        for v, c in zip(self.coeff_views(out), coeffs):
            v.copy_(c if isinstance(c, torch.Tensor) else torch.from_numpy(np.asarray(c)))
        return out

    def _check(self, x, total, name):
        torch = _torch()
        if not isinstance(x, torch.Tensor) or not x.is_cuda or x.dtype != torch.complex64:
            raise TypeError(f"{name}: expected a complex64 CUDA tensor")
        if not x.is_contiguous() or x.numel() != total:
            raise ValueError(f"{name}: expected a contiguous packed array of {total} elements, got {x.numel()}")

    def _stream(self):
        return C.c_void_p(_torch().cuda.current_stream(self.device).cuda_stream)

    # ---- transforms
    def isoft(self, coeffs, out=None):
        """Packed coefficients -> packed SO(3) functions."""
        self._check(coeffs, self.coeff_total, "coeffs")
        out = self.empty_grids() if out is None else out
        self._check(out, self.grid_total, "out")
        with _torch().cuda.device(self.device):
            _lib.check(_lib.lib().so3ft_group_isoft_c64(self._handle, C.c_void_p(coeffs.data_ptr()),
                                                        C.c_void_p(out.data_ptr()), self._stream()))
        return out

    def soft(self, grids, out=None):
        """Packed SO(3) functions -> packed coefficients."""
        self._check(grids, self.grid_total, "grids")
        out = self.empty_coeffs() if out is None else out
        self._check(out, self.coeff_total, "out")
        with _torch().cuda.device(self.device):
            _lib.check(_lib.lib().so3ft_group_soft_c64(self._handle, C.c_void_p(grids.data_ptr()),
                                                       C.c_void_p(out.data_ptr()), self._stream()))
        return out

    # ---- launch shapes
    def autotune(self):
        with _torch().cuda.device(self.device):
            _lib.check(_lib.lib().so3ft_group_autotune(self._handle, self._stream()))
        return self.config

    @property
    def config(self):
        """dict(inv_tb, inv_grid, fwd_tb, fwd_nw, fwd_grid, tuned, inv_sched, fwd_sched)."""
        c = np.zeros(8, np.int32)
        _lib.check(_lib.lib().so3ft_group_get_config(self._handle, c.ctypes.data))
        return dict(zip(("inv_tb", "inv_grid", "fwd_tb", "fwd_nw", "fwd_grid", "tuned", "inv_sched", "fwd_sched"),
                        map(int, c)))

    @config.setter
    def config(self, cfg):
        c = np.array([cfg["inv_tb"], cfg["inv_grid"], cfg["fwd_tb"], cfg["fwd_nw"], cfg["fwd_grid"], 0,
                      cfg.get("inv_sched", -1), cfg.get("fwd_sched", -1)], np.int32)
        _lib.check(_lib.lib().so3ft_group_set_config(self._handle, c.ctypes.data))

    def set_graphs(self, enable):
        _lib.check(_lib.lib().so3ft_group_set_graphs(self._handle, 1 if enable else 0))
