"""Host-side index helpers with the semantics of the reference's ``pysofft.utils``
(the f2py view of pysofft/utils.f90).  Functions that the reference returns as
1-based Fortran indices keep that convention, so ``soft.py``-style callers that
subtract 1 work unchanged.
"""
from functools import lru_cache

import numpy as np


def total_num_coeffs(bw):
    """utils.f90:32-36: sum_{l<bw} (2l+1)^2 = (4bw^3-bw)/3."""
    bw = int(bw)
    return (4 * bw ** 3 - bw) // 3


def euler_shape(bw):
    """utils.f90:364-368."""
    return np.array([2 * int(bw)] * 3, dtype=np.int64)


def n_lmc(bw):
    return int(bw) ** 2


def n_lmr(bw):
    bw = int(bw)
    return bw * (bw + 1) // 2


def lmc(l, m):
    """utils.f90:735-741 (1-based): l(l+1)+m+1."""
    return int(l) * (int(l) + 1) + int(m) + 1


def lmc_slice(l):
    """utils.f90:743-748 (1-based, inclusive)."""
    return np.array([lmc(l, -l), lmc(l, l)], dtype=np.int64)


def mlr(m, l, bw):
    """utils.f90:758-764 (1-based)."""
    m, l, bw = int(m), int(l), int(bw)
    return (m * (2 * bw - 1 - m)) // 2 + l + 1


def mlc(m, l, bw):
    """utils.f90:781-787 (1-based)."""
    m, l, bw = int(m), int(l), int(bw)
    return max(0, abs(m) * (2 * bw - abs(m)) - bw) + (0 if m >= 0 else bw - abs(m)) + l + 1


def triangular_to_flat_index(i, j, bw):
    """utils.f90:652-656 (1-based)."""
    return (i * (2 * bw - i + 1)) // 2 + j - i + 1


@lru_cache(maxsize=64)
def _slice_table(bw):
    """0-based start of f^l_{m,n} (l = max(|m|,|n|)..bw-1), indexed [m+bw-1, n+bw-1].

    The mnl layout is the storage order of utils.f90:140-201: for 0<=m1<=m2<bw the
    variants (m1,m2),(-m2,-m1),(m2,m1),(-m1,-m2),(m1,-m2),(-m1,m2),(m2,-m1),(-m2,m1)
    (duplicates skipped) each own bw-m2 consecutive entries.
    """
    W = 2 * bw - 1
    t = np.zeros((W, W), dtype=np.int64)
    pos = 0
    for m1 in range(bw):
        for m2 in range(m1, bw):
            L = bw - m2
            order = [(m1, m2)]
            if not (m1 == 0 and m2 == 0):
                order.append((-m2, -m1))
                if m1 != m2:
                    order += [(m2, m1), (-m1, -m2)]
                if m1 != 0 and m2 != 0:
                    order += [(m1, -m2), (-m1, m2)]
                    if m1 != m2:
                        order += [(m2, -m1), (-m2, m1)]
            for (m, n) in order:
                t[m + bw - 1, n + bw - 1] = pos
                pos += L
    t.flags.writeable = False
    return t


def coeff_slice_mnl(m, n, bw):
    """utils.f90:87-116 (1-based, inclusive)."""
    bw = int(bw)
    s = int(_slice_table(bw)[int(m) + bw - 1, int(n) + bw - 1])
    a = max(abs(int(m)), abs(int(n)))
    return np.array([s + 1, s + bw - a], dtype=np.int64)


def coeff_location_mnl(m, n, l, bw):
    """utils.f90:122-133 (1-based)."""
    bw = int(bw)
    a = max(abs(int(m)), abs(int(n)))
    if l < a or l >= bw:
        raise IndexError(f"l={l} does not satisfy max(|m|,|n|)={a} <= l < bw={bw}")
    return int(_slice_table(bw)[int(m) + bw - 1, int(n) + bw - 1]) + int(l) - a + 1


@lru_cache(maxsize=32)
def _coeff_degrees(bw):
    out = np.zeros((total_num_coeffs(bw), 3), dtype=np.int64)
    t = _slice_table(bw)
    for m in range(-bw + 1, bw):
        for n in range(-bw + 1, bw):
            a = max(abs(m), abs(n))
            s = t[m + bw - 1, n + bw - 1]
            ls = np.arange(a, bw)
            out[s:s + len(ls), 0] = ls
            out[s:s + len(ls), 1] = m
            out[s:s + len(ls), 2] = n
    out.flags.writeable = False
    return out


def get_coeff_degrees(bw):
    """utils.f90:140-201: (l, m, n) of every coefficient in storage order, shape (NC, 3)."""
    return _coeff_degrees(int(bw)).copy()


def legendre_quadrature_weights(bw):
    """utils.f90:389-404: Driscoll-Healy weights of the 2bw Chebyshev nodes (sum 2)."""
    bw = int(bw)
    xi = np.pi / (4 * bw)
    k = np.arange(2 * bw)[:, None]
    j = np.arange(bw)[None, :]
    s = np.sum((1.0 / (2 * j + 1)) * np.sin((2 * k + 1) * (2 * j + 1) * xi), axis=1)
    return s * (2.0 / bw) * np.sin((2 * np.arange(2 * bw) + 1) * xi)


def enforce_real_sym_mnl(coeff, bw):
    """utils.f90:471-488 on a host array, in place: f^l_{-m,-n} = (-1)^{m+n} conj f^l_{m,n}.

    The reference's loop order decides which half is the source: for m > 0 the
    m > 0 side, for m = 0 the n < 0 side; f^l_{0,0} is made real.
    """
    bw = int(bw)
    c = coeff
    t = _slice_table(bw)
    s00 = t[bw - 1, bw - 1]
    c[s00:s00 + bw] = c[s00:s00 + bw].real
    for m in range(0, bw):
        for n in range(-bw + 1, bw):
            if m == 0 and n <= 0:
                continue
            a = max(m, abs(n))
            L = bw - a
            src = t[m + bw - 1, n + bw - 1] if m > 0 else t[bw - 1, -n + bw - 1]
            dst = t[-m + bw - 1, -n + bw - 1] if m > 0 else t[bw - 1, n + bw - 1]
            sg = -1.0 if (m + n) % 2 else 1.0
            c[dst:dst + L] = sg * np.conj(c[src:src + L])
    return c


def enforce_real_sym_mnl_many(coeffs, bw):
    """utils.f90:497-505: coeffs is (NC, n) as the f2py view; also accepts (n, NC) C arrays."""
    arr = coeffs
    if arr.ndim == 2 and arr.shape[0] == total_num_coeffs(bw) and arr.shape[1] != arr.shape[0]:
        for i in range(arr.shape[1]):
            col = np.ascontiguousarray(arr[:, i])
            enforce_real_sym_mnl(col, bw)
            arr[:, i] = col
    else:
        for i in range(arr.shape[0]):
            enforce_real_sym_mnl(arr[i], bw)
    return arr
