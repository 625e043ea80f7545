"""Brute-force restatements from the closed-form definitions -- TEST INFRASTRUCTURE ONLY.

These are independent of both the C oracle and the reference code: they evaluate
the Wigner functions from the Jacobi-polynomial formula of docs/defs/wigner.md:9-16
and sum the transform directly.  Small sizes only (O(bw^6)).
"""
from math import lgamma

import numpy as np
from scipy.special import eval_jacobi


def wigner_d(l, m, n, beta):
    """Unnormalised small d^l_{m,n}(beta), docs/defs/wigner.md:11-14 (A_l = 1)."""
    beta = np.asarray(beta, dtype=np.float64)
    mu, nu = abs(m - n), abs(m + n)
    s = l - (mu + nu) // 2
    if s < 0:
        return np.zeros_like(beta)
    sign = -1.0 if ((m - n + mu) // 2) % 2 else 1.0
    lognorm = 0.5 * (lgamma(s + 1) + lgamma(s + mu + nu + 1) - lgamma(s + mu + 1) - lgamma(s + nu + 1))
    return (sign * np.exp(lognorm) * np.sin(beta / 2) ** mu * np.cos(beta / 2) ** nu
            * eval_jacobi(s, mu, nu, np.cos(beta)))


def wigner_d_norm(l, m, n, beta):
    """Normalised d~ = sqrt((2l+1)/2) d (docs/defs/wigner.md:14)."""
    return np.sqrt((2 * l + 1) / 2.0) * wigner_d(l, m, n, beta)


def grid_angles(bw):
    """(alpha, beta, gamma) samples, docs/defs/so3_grid.md:21-26."""
    N = 2 * bw
    a = np.arange(N) * np.pi / bw
    b = (np.arange(N) + 0.5) * np.pi / (2 * bw)
    return a, b, a.copy()


def basis(l, m, n, bw):
    """isoft basis function on the grid, C-order (gamma, alpha, beta):
    e^{-i m alpha} d~^l_{mn}(beta) e^{-i n gamma} / (2 pi)."""
    a, b, g = grid_angles(bw)
    d = wigner_d_norm(l, m, n, b)
    return (np.exp(-1j * n * g)[:, None, None] * np.exp(-1j * m * a)[None, :, None]
            * d[None, None, :]) / (2 * np.pi)


def isoft_direct(coeff, lmn, bw):
    """Direct sum f = sum c_lmn basis_lmn over the grid (lmn = coeff degree table)."""
    out = np.zeros((2 * bw,) * 3, np.complex128)
    a, b, g = grid_angles(bw)
    ea = np.exp(-1j * np.outer(np.arange(-bw + 1, bw), a))  # [m][alpha]
    eg = np.exp(-1j * np.outer(np.arange(-bw + 1, bw), g))
    cache = {}
    for idx, (l, m, n) in enumerate(lmn):
        c = coeff[idx]
        if c == 0:
            continue
        key = (l, m, n)
        if key not in cache:
            cache[key] = wigner_d_norm(l, m, n, b)
        d = cache[key]
        out += c * (eg[n + bw - 1][:, None, None] * ea[m + bw - 1][None, :, None] * d[None, None, :])
    return out / (2 * np.pi)


def soft_direct(f, lmn, bw, weights):
    """Direct quadrature f^l_{mn} = (pi/bw)^2 sum w_k f conj(e^{-im a} d~ e^{-in g}) / (2pi) * (2pi)^2 / (2pi)...

    Written as the exact inverse of isoft_direct on band-limited input:
    f^l_{mn} = (2 pi / (2bw)^2) sum_{j,j',k} w_k f(g_j, a_j', b_k) e^{+i m a} e^{+i n g} d~^l_{mn}(b_k).
    """
    a, b, g = grid_angles(bw)
    N = 2 * bw
    out = np.zeros(len(lmn), np.complex128)
    ea = np.exp(1j * np.outer(np.arange(-bw + 1, bw), a))
    eg = np.exp(1j * np.outer(np.arange(-bw + 1, bw), g))
    for idx, (l, m, n) in enumerate(lmn):
        d = wigner_d_norm(l, m, n, b)
        proj = np.einsum('gab,g,a->b', f, eg[n + bw - 1], ea[m + bw - 1])
        out[idx] = (2 * np.pi / N ** 2) * np.sum(weights * proj * d)
    return out


def rotate_ylm_direct(ylm, euler, bw):
    """soft.py:713-764 restated with Jacobi d: y'^l_n = sum_m y^l_m d^l_{n,m}(beta) e^{-im a} e^{-in g}."""
    al, be, ga = euler
    out = np.zeros(bw * bw, np.complex128)
    for l in range(bw):
        for n in range(-l, l + 1):
            acc = 0j
            for m in range(-l, l + 1):
                acc += ylm[l * (l + 1) + m] * wigner_d(l, n, m, be) * np.exp(-1j * m * al)
            out[l * (l + 1) + n] = acc * np.exp(-1j * n * ga)
    return out


def dlml_closed_form(l, m, beta):
    """d~^l_{m,l}(beta) closed form, tests/test_fortran_backend.py:8-12."""
    from math import factorial as f
    return (np.sqrt((2 * l + 1) / 2) * np.sqrt(f(2 * l) / (f(l + m) * f(l - m)))
            * np.cos(beta / 2) ** (l + m) * np.sin(beta / 2) ** (l - m))


def cos_2eta_coeff(l, eta):
    """f^l_{00} of (2 eta+1) cos(beta)^(2 eta), tests/test_fortran_backend.py:104-111."""
    if l <= 2 * eta and l % 2 == 0:
        kcos = (2 * eta + 1) * np.sqrt((2 * l + 1) / 2) * 2 * np.pi
        return kcos * np.exp(lgamma(2 * eta + 1) + lgamma(eta + l // 2 + 1) + (l + 1) * np.log(2)
                             - lgamma(eta - l // 2 + 1) - lgamma(2 * eta + l + 2))
    return 0.0


def cos_2eta_func(beta, eta):
    return (2 * eta + 1) * np.cos(beta) ** (2 * eta)
