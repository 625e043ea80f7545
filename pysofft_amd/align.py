"""Learned alignment proposal with local correlation refinement.

Pipeline for pairs (f, g) of spherical-harmonic coefficient sets (LMc order, bw <= 64)
with g ~ D(R) f:
  1. features: the l <= 8 coefficients of f and g and the per-degree powers of both
     (``so3ft_align_features``);
  2. proposal: a small MLP (BF16 MFMA, FP32 accumulation; ``so3ft_mlp_*``) maps the
     features to a 6-D rotation representation -> rotation R0 -> ZYZ Euler angles;
  3. refinement: the FP64 correlation |C| on 5 x 5 x 5 windows of successively finer
     grids (bandwidth 4, 8, 16, 32, 128 by default, +-2 steps each, degrees below the
     grid bandwidth up to 64) around the estimate
     (``so3ft_refine_correlation_peaks``);
  4. fallback: pairs whose refined peak, normalised by ||f|| ||g||, stays below a
     threshold are aligned by the full FP64 correlation at bw and its peak search
     (``Soft.correlation_peaks``, the method of the batched correlation workload), then the
     last refinement level around that peak.  ``fallback="coarse"`` selects a cheaper
     exhaustive search instead: the full correlation of the degrees below bw/2 (a grid of
     (bw)^3 points, 8x cheaper), then windows at grid bandwidths bw and 2 bw.

``proposer="grid"`` replaces the network by the argmax of the full correlation of the
degrees below 4 (an 8^3 grid): the same refinement ladder and fallback without the
learned part, the ablation that measures what the network contributes.

The network is trained on the GPU by the library's own forward/backward/Adam kernels on
pairs generated with exact Wigner-D rotations of the l <= 8 part (``training_batch``).
No counterpart in the reference; the correlation it refines is the reference's
(pysofft/soft.f08:2242-2420).
"""
import ctypes as C
import math

import numpy as np

from . import _lib, rotate as _rotate
from . import workloads as W

N_FEAT = 452
DIMS = (N_FEAT, 512, 512, 512, 6)
LEVELS = ((4, 4), (8, 8), (16, 16), (32, 32), (128, 64))   # (grid bandwidth, degree limit)


def _torch():
    import torch
    return torch


def _stream(t):
    torch = _torch()
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _mcheck(rc):
    if rc != 0:
        msg = _lib.lib().so3ft_mlp_last_error()
        raise _lib.So3ftError(f"so3ft error {rc}: {msg.decode() if msg else ''}")


def features(f_lms, g_lms, mean=None, std=None):
    """(n, bw^2) complex128 CUDA tensors -> (n, 452) float32 features on the device."""
    torch = _torch()
    f = f_lms.to(torch.complex128).contiguous()
    g = g_lms.to(torch.complex128).contiguous()
    n, nb = f.shape
    bw = int(round(math.sqrt(nb)))
    x = torch.empty((n, N_FEAT), dtype=torch.float32, device=f.device)
    mp = C.c_void_p(mean.data_ptr()) if mean is not None else None
    sp = C.c_void_p(std.data_ptr()) if std is not None else None
    _mcheck(_lib.lib().so3ft_align_features(bw, C.c_void_p(f.data_ptr()), C.c_void_p(g.data_ptr()), n, mp, sp,
                                            C.c_void_p(x.data_ptr()), _stream(f)))
    return x


def init_params(dims=DIMS, seed=0):
    """He-uniform weights, zero biases, float32 host vector in the network's layout."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(len(dims) - 1):
        fan_in, fan_out = dims[i], dims[i + 1]
        lim = math.sqrt(6.0 / fan_in) if i + 2 < len(dims) else math.sqrt(1.0 / fan_in)
        out.append(rng.uniform(-lim, lim, (fan_out, fan_in)).astype(np.float32).ravel())
        out.append(np.zeros(fan_out, np.float32))
    return np.concatenate(out)


class AlignmentMLP:
    """Handle on the device network (weights, Adam state, activations)."""

    def __init__(self, dims=DIMS, seed=0, params=None):
        self.dims = tuple(int(d) for d in dims)
        arr = (C.c_int32 * len(self.dims))(*self.dims)
        h = C.c_int64()
        _mcheck(_lib.lib().so3ft_mlp_create(arr, len(self.dims) - 1, C.byref(h)))
        self._h = h.value
        self.set_params(init_params(self.dims, seed) if params is None else params)
        self.mean = None
        self.std = None

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            try:
                _lib.lib().so3ft_mlp_destroy(self._h)
            except Exception:
                pass
            self._h = None

    @property
    def num_params(self):
        n = C.c_int64()
        _mcheck(_lib.lib().so3ft_mlp_num_params(self._h, C.byref(n)))
        return n.value

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        if p.size != self.num_params:
            raise ValueError(f"expected {self.num_params} parameters, got {p.size}")
        _mcheck(_lib.lib().so3ft_mlp_set_params(self._h, p.ctypes.data_as(C.c_void_p)))

    def get_params(self, grads=False):
        p = np.empty(self.num_params, np.float32)
        _mcheck(_lib.lib().so3ft_mlp_get_params(self._h, int(grads), p.ctypes.data_as(C.c_void_p)))
        return p

    def forward(self, x):
        """x (n, dims[0]) float32 CUDA -> (R (n, 3, 3), raw outputs (n, 6))."""
        torch = _torch()
        x = x.to(torch.float32).contiguous()
        n = x.shape[0]
        R = torch.empty((n, 9), dtype=torch.float32, device=x.device)
        o = torch.empty((n, 6), dtype=torch.float32, device=x.device)
        _mcheck(_lib.lib().so3ft_mlp_forward(self._h, C.c_void_p(x.data_ptr()), n, C.c_void_p(R.data_ptr()),
                                             C.c_void_p(o.data_ptr()), _stream(x)))
        return R.view(n, 3, 3), o

    def train_step(self, x, T, lr):
        """One forward/backward (+ Adam when lr > 0); T (n, 3, 3) target rotations.
        Returns the mean loss ||R - T||_F^2."""
        torch = _torch()
        x = x.to(torch.float32).contiguous()
        T = T.reshape(-1, 9).to(torch.float32).contiguous()
        loss = C.c_float()
        _mcheck(_lib.lib().so3ft_mlp_train_step(self._h, C.c_void_p(x.data_ptr()), C.c_void_p(T.data_ptr()),
                                                x.shape[0], float(lr), C.byref(loss), _stream(x)))
        return loss.value


def _quat_to_matrix(q):
    torch = _torch()
    w, x, y, z = q.unbind(1)
    R = torch.stack([1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                     2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                     2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)], 1)
    return R.view(-1, 3, 3)


def _matrix_to_euler_t(R):
    torch = _torch()
    beta = torch.arccos(R[:, 2, 2].clamp(-1.0, 1.0))
    alpha = torch.atan2(R[:, 1, 2], R[:, 0, 2])
    gamma = torch.atan2(R[:, 2, 1], -R[:, 2, 0])
    small = torch.sin(beta).abs() < 1e-12
    alpha = torch.where(small, torch.atan2(R[:, 1, 0], R[:, 0, 0]), alpha)
    gamma = torch.where(small, torch.zeros_like(gamma), gamma)
    two_pi = 2 * math.pi
    return torch.stack([torch.remainder(alpha, two_pi), beta, torch.remainder(gamma, two_pi)], 1)


def _euler_to_matrix_t(e):
    torch = _torch()
    a, b, g = e[:, 0], e[:, 1], e[:, 2]
    ca, sa, cb, sb, cg, sg = torch.cos(a), torch.sin(a), torch.cos(b), torch.sin(b), torch.cos(g), torch.sin(g)
    # Rz(a) Ry(b) Rz(g)
    R = torch.stack([ca * cb * cg - sa * sg, -ca * cb * sg - sa * cg, ca * sb,
                     sa * cb * cg + ca * sg, -sa * cb * sg + ca * cg, sa * sb,
                     -sb * cg, sb * sg, cb], 1)
    return R.view(-1, 3, 3)


def training_batch(n, gen, bw=64, snr=10.0, device="cuda"):
    """n synthetic pairs for training, on the device: (features-ready f, g arrays, target R).

    f: iid complex N(0,1) with power ~ (1+l)^-1 for l < bw; a Haar-uniform rotation R (ZYZ
    angles e); g's l <= 8 part = D(e) f_{l<=8} (exact Wigner-D, Risbo recurrence) plus
    complex Gaussian noise at `snr`; g's l > 8 part = f_l + noise', which has exactly the
    distribution of (D(e) f + noise)_l's per-degree power (noise is isotropic), the only
    l > 8 quantity the features use."""
    torch = _torch()
    nb = bw * bw
    l = torch.floor(torch.sqrt(torch.arange(nb, device=device, dtype=torch.float64)))
    scale = (1.0 + l) ** -0.5
    f = torch.complex(torch.randn((n, nb), generator=gen, dtype=torch.float64, device=device),
                      torch.randn((n, nb), generator=gen, dtype=torch.float64, device=device)) * scale
    q = torch.randn((n, 4), generator=gen, dtype=torch.float64, device=device)
    q = q / q.norm(dim=1, keepdim=True)
    R = _quat_to_matrix(q)
    e = _matrix_to_euler_t(R)
    noise = torch.complex(torch.randn((n, nb), generator=gen, dtype=torch.float64, device=device),
                          torch.randn((n, nb), generator=gen, dtype=torch.float64, device=device))
    p_sig = (f.abs() ** 2).sum(1, keepdim=True)
    p_noise = (noise.abs() ** 2).sum(1, keepdim=True)
    noise = noise * torch.sqrt(p_sig / (snr * p_noise))
    glow = _rotate.rotate_ylm_device(f[:, :81].contiguous(), e.contiguous(), paired=True)
    g = f + noise
    g[:, :81] = glow + noise[:, :81]
    return f, g, R.to(torch.float32), e


class Aligner:
    """Learned proposal + local refinement + full-correlation fallback."""

    def __init__(self, mlp=None, levels=LEVELS, threshold=0.6, device="cuda", fallback="full", proposer="mlp"):
        if fallback not in ("full", "coarse"):
            raise ValueError("fallback must be 'full' or 'coarse'")
        if proposer not in ("mlp", "grid"):
            raise ValueError("proposer must be 'mlp' or 'grid'")
        self.mlp = mlp if mlp is not None else (AlignmentMLP() if proposer == "mlp" else None)
        self.levels = tuple(levels)
        self.threshold = threshold
        self.device = device
        self.fallback = fallback
        self.proposer = proposer
        self._soft = {}

    def _plan(self, bw):
        from .soft import Soft
        if bw not in self._soft:
            self._soft[bw] = Soft(bw)
        return self._soft[bw]

    @staticmethod
    def _peak_euler(pk, bw):
        """grid indices (i_alpha, i_beta, i_gamma) of a bandwidth-bw grid -> ZYZ angles"""
        torch = _torch()
        a = pk[:, 0].to(torch.float64) * (math.pi / bw)
        b = (2 * pk[:, 1].to(torch.float64) + 1) * (math.pi / (4 * bw))
        c = pk[:, 2].to(torch.float64) * (math.pi / bw)
        return torch.stack([a, b, c], 1).contiguous()

    # ---- training -----------------------------------------------------------------
    def fit_normalisation(self, n=65536, seed=1, batch=8192):
        torch = _torch()
        gen = torch.Generator(device=self.device).manual_seed(seed)
        xs = []
        for i in range(0, n, batch):
            f, g, _, _ = training_batch(min(batch, n - i), gen, device=self.device)
            xs.append(features(f, g))
        x = torch.cat(xs)
        self.mlp.mean = x.mean(0).contiguous()
        self.mlp.std = (x.std(0) + 1e-6).contiguous()

    def train(self, pairs=24_576_000, batch=4096, lr=3e-3, seed=2, log_every=0):
        """Train on `pairs` freshly generated pairs (one step per batch), cosine-decayed lr."""
        torch = _torch()
        if self.mlp.mean is None:
            self.fit_normalisation()
        gen = torch.Generator(device=self.device).manual_seed(seed)
        steps = max(1, pairs // batch)
        hist = []
        for k in range(steps):
            f, g, T, _ = training_batch(batch, gen, device=self.device)
            x = features(f, g, self.mlp.mean, self.mlp.std)
            lk = lr * 0.5 * (1.0 + math.cos(math.pi * k / steps))
            loss = self.mlp.train_step(x, T, lk)
            hist.append(loss)
            if log_every and k % log_every == 0:
                print(f"step {k}/{steps} loss {loss:.4f}", flush=True)
        return hist

    # ---- inference ----------------------------------------------------------------
    def propose(self, f_lms, g_lms):
        """(n, bw^2) pairs -> proposed ZYZ Euler angles (n, 3) float64 on the device: the
        network's rotation, or with proposer="grid" the peak of the degrees-below-4
        correlation."""
        torch = _torch()
        if self.proposer == "grid":
            f = f_lms.to(torch.complex128)[:, :16].contiguous()
            g = g_lms.to(torch.complex128)[:, :16].contiguous()
            pk, _ = self._plan(4).correlation_peaks(f, g)
            return self._peak_euler(pk, 4)
        x = features(f_lms, g_lms, self.mlp.mean, self.mlp.std)
        R, _ = self.mlp.forward(x)
        return _matrix_to_euler_t(R.to(torch.float64)).contiguous()

    def refine(self, f_lms, g_lms, euler, levels=None):
        """Window refinement through `levels`; returns (euler (n, 3), |C| of the last level)."""
        torch = _torch()
        f = f_lms.to(torch.complex128).contiguous()
        g = g_lms.to(torch.complex128).contiguous()
        n, nb = f.shape
        bw = int(round(math.sqrt(nb)))
        e = euler.to(torch.float64).contiguous()
        score = torch.empty(n, dtype=torch.float64, device=f.device)
        for grid_bw, lmax in (levels or self.levels):
            out = torch.empty_like(e)
            _mcheck(_lib.lib().so3ft_refine_correlation_peaks(
                bw, C.c_void_p(f.data_ptr()), C.c_void_p(g.data_ptr()), n, C.c_void_p(e.data_ptr()),
                C.c_void_p(out.data_ptr()), C.c_void_p(score.data_ptr()), int(grid_bw), int(min(lmax, bw)), _stream(f)))
            e = out
        return e, score

    def align(self, f_lms, g_lms):
        """Rotations (ZYZ Euler, (n, 3)) aligning g to f, and the mask of pairs that fell back
        to the exhaustive search."""
        torch = _torch()
        f = f_lms.to(torch.complex128).contiguous()
        g = g_lms.to(torch.complex128).contiguous()
        n, nb = f.shape
        bw = int(round(math.sqrt(nb)))
        e0 = self.propose(f, g)
        e, score = self.refine(f, g, e0)
        norm = torch.linalg.vector_norm(f, dim=1) * torch.linalg.vector_norm(g, dim=1)
        bad = score / norm < self.threshold
        if bool(bad.any()):
            idx = torch.nonzero(bad).flatten()
            fi, gi = f[idx].contiguous(), g[idx].contiguous()
            if self.fallback == "full":
                # the full FP64 correlation at bw and its peak (soft.f08:2242-2420), then
                # the last refinement level around that grid point
                pk, _ = self._plan(bw).correlation_peaks(fi, gi)
                ef, _ = self.refine(fi, gi, self._peak_euler(pk, bw), levels=self.levels[-1:])
            else:
                fb = max(bw // 2, 2)
                pk, _ = self._plan(fb).correlation_peaks(fi[:, :fb * fb].contiguous(), gi[:, :fb * fb].contiguous())
                # the coarse peak is within half a step (pi / (2 fb)) of the optimum: a window
                # of +-2 steps at grid bandwidth bw covers it, then one at 2 bw
                ef, _ = self.refine(fi, gi, self._peak_euler(pk, fb), levels=((bw, bw), (2 * bw, bw)))
            e = e.clone()
            e[idx] = ef
        return e, bad
