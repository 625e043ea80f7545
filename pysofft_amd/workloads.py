"""Synthetic workload generators for the alignment (rotational cross-correlation) path.

Signals follow the measurement definition of the batched correlation workload: f has
iid complex N(0,1) spherical-harmonic coefficients with power decaying as (1+l)^-1,
g is f rotated by a Haar-uniform rotation through the exact Wigner-D kernel, plus
complex Gaussian noise at a given SNR (signal power / noise power).
"""
import numpy as np

from . import rotate as _rotate


def haar_rotations(n, rng):
    """n Haar-uniform rotation matrices (uniform unit quaternions)."""
    q = rng.standard_normal((n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _rz(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(np.shape(t) + (3, 3))
    R[..., 0, 0] = c; R[..., 0, 1] = -s; R[..., 1, 0] = s; R[..., 1, 1] = c; R[..., 2, 2] = 1
    return R


def _ry(t):
    c, s = np.cos(t), np.sin(t)
    R = np.zeros(np.shape(t) + (3, 3))
    R[..., 0, 0] = c; R[..., 0, 2] = s; R[..., 2, 0] = -s; R[..., 2, 2] = c; R[..., 1, 1] = 1
    return R


def euler_to_matrix(e):
    """ZYZ Euler angles (..., 3) -> R = Rz(alpha) Ry(beta) Rz(gamma)."""
    e = np.asarray(e, dtype=np.float64)
    return _rz(e[..., 0]) @ _ry(e[..., 1]) @ _rz(e[..., 2])


def matrix_to_euler(R):
    """Inverse of euler_to_matrix: alpha, gamma in [0, 2pi), beta in [0, pi]."""
    R = np.asarray(R, dtype=np.float64)
    beta = np.arccos(np.clip(R[..., 2, 2], -1.0, 1.0))
    alpha = np.arctan2(R[..., 1, 2], R[..., 0, 2])
    gamma = np.arctan2(R[..., 2, 1], -R[..., 2, 0])
    small = np.abs(np.sin(beta)) < 1e-12
    # gimbal lock: only alpha + gamma (or alpha - gamma) is defined
    alpha = np.where(small, np.arctan2(R[..., 1, 0], R[..., 0, 0]), alpha)
    gamma = np.where(small, 0.0, gamma)
    return np.stack([np.mod(alpha, 2 * np.pi), beta, np.mod(gamma, 2 * np.pi)], axis=-1)


def angular_distance(R1, R2):
    """Geodesic angle between rotations (radians)."""
    t = np.einsum("...ij,...ij->...", R1, R2)
    return np.arccos(np.clip((t - 1.0) / 2.0, -1.0, 1.0))


def random_sh(n, bw, rng, decay=1.0):
    """(n, bw^2) LMc coefficients, iid complex N(0,1) scaled so power ~ (1+l)^-decay."""
    l = np.floor(np.sqrt(np.arange(bw * bw))).astype(np.int64)
    c = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    return c * (1.0 + l) ** (-decay / 2.0)


def rotated_pairs(n, bw, rng, snr=10.0, euler=None):
    """(f, g, euler): g = D(euler) f + complex Gaussian noise with ||D f||^2 / ||noise||^2 = snr
    per pair.  The rotation runs on the GPU (exact Wigner-D, Risbo recurrence)."""
    import torch
    f = random_sh(n, bw, rng)
    if euler is None:
        euler = matrix_to_euler(haar_rotations(n, rng))
    g = _rotate.rotate_ylm_device(torch.from_numpy(f).cuda(), torch.from_numpy(np.ascontiguousarray(euler)).cuda(),
                                  paired=True).cpu().numpy()
    if snr is not None and np.isfinite(snr):
        noise = rng.standard_normal(g.shape) + 1j * rng.standard_normal(g.shape)
        p_sig = np.sum(np.abs(g) ** 2, axis=1, keepdims=True)
        p_noise = np.sum(np.abs(noise) ** 2, axis=1, keepdims=True)
        g = g + noise * np.sqrt(p_sig / (snr * p_noise))
    return f, g, euler
