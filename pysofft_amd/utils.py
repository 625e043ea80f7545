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


# ---- lmn order (Risbo recurrence layout), utils.f90:228-355, 510-560 -----------------
# Degree-major blocks of sum_{l'<l} (2l'+1)^2 entries; inside a degree the pairs
# 0 <= m1 <= m2 <= l (m1 outer) each own their distinct variants in the order
# (m1,m2), (-m2,-m1), (m2,m1), (-m1,-m2), (m1,-m2), (-m1,m2), (m2,-m1), (-m2,m1):
# 1 entry for (0,0), 4 when m1 == 0 or m1 == m2, else 8.

def _pair_variants(m1, m2):
    v = [(m1, m2)]
    if m1 == 0 and m2 == 0:
        return v
    v.append((-m2, -m1))
    if m1 != m2:
        v += [(m2, m1), (-m1, -m2)]
    if m1 != 0:
        v += [(m1, -m2), (-m1, m2)]
        if m1 != m2:
            v += [(m2, -m1), (-m2, m1)]
    return v


def coeff_slice_lmn(l, n1, n2):
    """utils.f90:228-258 (1-based, inclusive): the entries of the pair (min, max) of
    (|n1|, |n2|) at degree l."""
    l = int(l)
    m, n = min(abs(int(n1)), abs(int(n2))), max(abs(int(n1)), abs(int(n2)))
    start = l * (4 * l * l - 1) // 3
    for a in range(0, m + 1):
        for b in range(a, l + 1):
            if (a, b) == (m, n):
                return np.array([start + 1, start + len(_pair_variants(a, b))], dtype=np.int64)
            start += len(_pair_variants(a, b))
    raise IndexError(f"no pair ({m}, {n}) at degree {l}")


def coeff_location_lmn(l, n1, n2):
    """utils.f90:261-292 (1-based)."""
    s = int(coeff_slice_lmn(l, n1, n2)[0])
    m, n = min(abs(int(n1)), abs(int(n2))), max(abs(int(n1)), abs(int(n2)))
    return s + _pair_variants(m, n).index((int(n1), int(n2)))


@lru_cache(maxsize=32)
def _coeff_degrees_risbo(bw):
    rows = []
    for l in range(bw):
        for m1 in range(l + 1):
            for m2 in range(m1, l + 1):
                rows += [(l, a, b) for (a, b) in _pair_variants(m1, m2)]
    out = np.array(rows, dtype=np.int64).reshape(-1, 3)
    out.flags.writeable = False
    return out


def get_coeff_degrees_risbo(bw):
    """utils.f90:298-355: (l, m, n) of every coefficient in lmn storage order, (NC, 3)."""
    return _coeff_degrees_risbo(int(bw)).copy()


@lru_cache(maxsize=32)
def lmn_to_mnl(bw):
    """perm[i] = mnl position of the coefficient stored at lmn position i (0-based)."""
    bw = int(bw)
    d = _coeff_degrees_risbo(bw)
    t = _slice_table(bw)
    a = np.maximum(np.abs(d[:, 1]), np.abs(d[:, 2]))
    p = t[d[:, 1] + bw - 1, d[:, 2] + bw - 1] + d[:, 0] - a
    p.flags.writeable = False
    return p


def enforce_real_sym_lmn(coeff, bw):
    """utils.f90:510-550 on a host array, in place: f^l_{-m,-n} = (-1)^{m+n} conj f^l_{m,n}
    with the sources m > 0 and, for m = 0, n > 0 (the opposite side from the mnl
    routine); f^l_{0,0} is made real."""
    bw = int(bw)
    d = _coeff_degrees_risbo(bw)
    l, m, n = d[:, 0], d[:, 1], d[:, 2]
    key = (l * (2 * bw) + (m + bw)) * (2 * bw) + (n + bw)
    inv = np.empty(key.max() + 1, dtype=np.int64)
    inv[key] = np.arange(len(d))
    tgt = (m < 0) | ((m == 0) & (n < 0))
    src = inv[(l * (2 * bw) + (-m + bw)) * (2 * bw) + (-n + bw)]
    sg = np.where((m + n) % 2 == 0, 1.0, -1.0)
    zero = (m == 0) & (n == 0)
    coeff[zero] = coeff[zero].real
    coeff[tgt] = sg[tgt] * np.conj(coeff[src[tgt]])
    return coeff


def enforce_real_sym_lmn_many(coeffs, bw):
    """utils.f90:552-560: coeffs is (NC, n) as the f2py view; also accepts (n, NC) C arrays."""
    arr = coeffs
    if arr.ndim == 2 and arr.shape[0] == total_num_coeffs(bw) and arr.shape[1] != arr.shape[0]:
        for i in range(arr.shape[1]):
            col = np.ascontiguousarray(arr[:, i])
            enforce_real_sym_lmn(col, bw)
            arr[:, i] = col
    else:
        for i in range(arr.shape[0]):
            enforce_real_sym_lmn(arr[i], bw)
    return arr
