"""Pins the CPU oracle (oracle/) against the reference's own code and mathematics.

Each oracle function is checked against at least one of: vectors produced by the
reference's Fortran (tests/golden/ref_vectors.npz, and the live build in
oracle/_ref when present), the reference's analytic known-answer tests
(tests/test_fortran_backend.py:8-114, 329-461), or closed forms / invariants.
"""
import numpy as np
import pytest

import oracle as O
from oracle import RefLib, brute
from conftest import rel_l2


@pytest.mark.parametrize("bw", [1, 2, 3, 5, 8, 13])
def test_layout_vs_reference_vectors(golden, bw):
    assert np.array_equal(O.coeff_degrees(bw), golden[f"degrees_{bw}"].astype(np.int64))


@pytest.mark.parametrize("bw", [1, 2, 3, 4, 7, 16, 32, 64, 128, 256])
def test_weights_vs_reference_vectors(golden, bw):
    w = O.legendre_weights(bw)
    np.testing.assert_allclose(w, golden[f"weights_{bw}"], rtol=0, atol=4e-16 * bw)
    assert abs(w.sum() - 2.0) < 1e-13          # sum of weights = int sin = 2
    # exactness on even Legendre polynomials P_m(cos b), m <= bw (utils.f90:378)
    b = O.beta_samples(2 * bw)
    for m in range(0, min(bw, 12)):
        pm = np.polynomial.legendre.legval(np.cos(b), [0] * m + [1])
        assert abs(np.dot(w, pm) - (2.0 if m == 0 else 0.0)) < 1e-12


def test_seeds_vs_reference_vectors(golden):
    s = O.seeds(9)
    np.testing.assert_array_equal(s, golden["seeds_9"])
    # closed form d~^l_{m,l} (tests/test_fortran_backend.py:8-12)
    b = O.beta_samples(18)
    for m in range(9):
        for l in range(m, 9):
            tri = (m * (2 * 9 - m + 1)) // 2 + l - m
            np.testing.assert_allclose(s[tri], brute.dlml_closed_form(l, m, b), rtol=1e-13, atol=1e-15)


@pytest.mark.parametrize("p,q,bw", [(0, 0, 16), (3, 7, 16), (5, 5, 16), (0, 15, 16), (32, 56, 129), (0, 0, 128),
                                    (17, 90, 128), (127, 127, 128)])
def test_genwig_vs_reference_vectors(golden, p, q, bw):
    g = O.genwig(p, q, bw)
    ref = golden[f"genwig_{p}_{q}_{bw}"]
    assert np.abs(g - ref).max() < 1e-12 * max(1.0, np.abs(ref).max())


def test_genwig_vs_jacobi():
    bw = 12
    b = O.beta_samples(2 * bw)
    for (p, q) in [(0, 0), (0, 5), (2, 3), (4, 4), (1, 11)]:
        g = O.genwig(p, q, bw)
        for l in range(q, bw):
            np.testing.assert_allclose(g[l - q], brute.wigner_d_norm(l, p, q, b), rtol=1e-11, atol=1e-13)


def test_legendre_and_assoc_legendre():
    # d~^l_{00} = sqrt((2l+1)/2) P_l(cos b) (tests/test_fortran_backend.py:205-231)
    bw = 129
    b = O.beta_samples(2 * bw)
    g = O.genwig(0, 0, bw)
    P = np.array([np.polynomial.legendre.legval(np.cos(b), [0] * l + [1]) for l in range(bw)])
    np.testing.assert_allclose(g, P * np.sqrt((2 * np.arange(bw)[:, None] + 1) / 2), rtol=1e-9, atol=1e-11)


def test_risbo_step_vs_reference_vectors(golden):
    cb, sb = golden["risbo_cs"]
    sq = np.sqrt(np.arange(25))
    d = None
    for l in range(12):
        d = O.risbo_step(d, l, cb, sb, sq)
    np.testing.assert_array_equal(d, golden["risbo_l11"])


def test_risbo_reduced_tables_vs_reference_vectors(golden):
    """The symmetry-reduced Risbo recurrence (wigner.f90:626-748) through the mnl d table of
    genwig_all_risbo (wigner.f90:784-822), normalized and not, and the Kostelec table of
    genwig_all_kostelec (wigner.f90:423-465): bit-equal to the reference build's output."""
    np.testing.assert_array_equal(O.genwig_all_risbo(9, True), golden["genwig_all_risbo_9"])
    np.testing.assert_array_equal(O.genwig_all_risbo(5, False), golden["genwig_all_risbo_raw_5"])
    np.testing.assert_array_equal(O.genwig_all_kostelec(9), golden["genwig_all_kostelec_9"])


def _dtab_index(bw):
    """(p, q, l) of every row of the mnl d table (utils.f90:683-695 order)"""
    return [(p, q, l) for p in range(bw) for q in range(p, bw) for l in range(q, bw)]


def test_risbo_table_vs_jacobi_and_kostelec():
    """Independent of the reference: every entry of the Risbo table equals the normalised
    Jacobi-polynomial d~^l_{p,q}(beta_k) (docs/defs/wigner.md:11-14) and the Kostelec table
    to 1e-13; the unnormalised table is the normalised one / sqrt((2l+1)/2)."""
    bw = 7
    t = O.genwig_all_risbo(bw, True)
    b = O.beta_samples(2 * bw)
    for i, (p, q, l) in enumerate(_dtab_index(bw)):
        np.testing.assert_allclose(t[i], brute.wigner_d_norm(l, p, q, b), rtol=0, atol=1e-13)
    np.testing.assert_allclose(t, O.genwig_all_kostelec(bw), rtol=0, atol=1e-13)
    raw = O.genwig_all_risbo(bw, False)
    ls = np.array([l for (_, _, l) in _dtab_index(bw)], float)
    np.testing.assert_allclose(raw * np.sqrt((2 * ls + 1) / 2)[:, None], t, rtol=0, atol=1e-14)


def test_full_big_d_vs_reference_vectors_and_closed_form(golden):
    """full_big_d_risbo (wigner.f90:844-934): equal to the reference build's vectors, and to
    the closed form e^{-i m alpha} d^l_{mn}(beta) e^{-i n gamma} (/(2 pi) normalised) with the
    routine's own convention: a factor -1 on the entries with m n < 0 and m + n odd."""
    al, be, ga = golden["full_big_d_angles"]
    x6 = O.full_big_d_risbo(6, al, be, ga, True)
    np.testing.assert_allclose(x6, golden["full_big_d_6"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(O.full_big_d_risbo(4, al, be, ga, False), golden["full_big_d_raw_4"], rtol=0,
                               atol=1e-15)
    deg = O.coeff_degrees(6)
    ref = np.array([np.exp(-1j * m * al) * brute.wigner_d_norm(l, m, n, np.array([be]))[0] * np.exp(-1j * n * ga)
                    for l, m, n in deg]) / (2 * np.pi)
    flip = (deg[:, 1] * deg[:, 2] < 0) & ((deg[:, 1] + deg[:, 2]) % 2 == 1)
    ref[flip] *= -1
    np.testing.assert_allclose(x6, ref, rtol=0, atol=1e-14)


@pytest.mark.parametrize("bw", [1, 4, 6])
def test_enforce_real_sym_vs_reference_vectors(golden, bw):
    c = golden[f"sym_in_{bw}"].copy()
    O.enforce_real_sym(c, bw)
    np.testing.assert_array_equal(c, golden[f"sym_out_{bw}"])


@pytest.mark.parametrize("bw", [2, 3, 4, 8])
def test_isoft_vs_reference_dmatrices(golden, bw):
    P = O.OraclePlan(bw)
    f = P.isoft(golden[f"isoft_c_{bw}"])
    assert rel_l2(f, golden[f"isoft_f_{bw}"]) < 1e-13


@pytest.mark.parametrize("bw", [2, 3, 5])
def test_isoft_vs_jacobi_direct_sum(bw):
    rng = np.random.default_rng(bw)
    P = O.OraclePlan(bw)
    c = rng.standard_normal(P.nc) + 1j * rng.standard_normal(P.nc)
    assert rel_l2(P.isoft(c), brute.isoft_direct(c, O.coeff_degrees(bw), bw)) < 1e-13


def test_isoft_soft_kats():
    """tests/test_fortran_backend.py:329-461 at bw=32: unit coefficient -> D^l_mn on the grid."""
    bw = 32
    P = O.OraclePlan(bw)
    tab = O.slice_table(bw)
    for (l, m, n) in [(0, 0, 0), (1, 1, -1), (1, 0, 1), (2, -2, 1), (2, 1, -2), (2, 0, 0), (5, 3, 5), (9, -7, 9)]:
        c = np.zeros(P.nc, np.complex128)
        c[tab[m + bw - 1, n + bw - 1] + l - max(abs(m), abs(n))] = 1
        f = P.isoft(c)
        D = brute.basis(l, m, n, bw)
        assert np.abs(f - D).max() < 1e-13
        c2 = P.soft(D)
        assert np.abs(c2 - c).max() < 1e-12
    b = O.beta_samples(2 * bw)
    for eta in [1, 5, 15]:
        c = np.zeros(P.nc, np.complex128)
        for l in range(bw):
            c[tab[bw - 1, bw - 1] + l] = brute.cos_2eta_coeff(l, eta)
        f = P.isoft(c)
        np.testing.assert_allclose(f[0, 0, :], brute.cos_2eta_func(b, eta), rtol=1e-10, atol=1e-10)
        F = np.broadcast_to(brute.cos_2eta_func(b, eta)[None, None, :], f.shape).astype(np.complex128)
        np.testing.assert_allclose(P.soft(F), c, rtol=1e-10, atol=1e-10)


@pytest.mark.parametrize("bw", [1, 2, 3, 7, 10])
def test_invertibility_and_real(bw):
    rng = np.random.default_rng(100 + bw)
    P = O.OraclePlan(bw)
    c = rng.standard_normal(P.nc) + 1j * rng.standard_normal(P.nc)
    f = P.isoft(c)
    assert rel_l2(P.soft(f), c) < 1e-13
    cs = c.copy()
    O.enforce_real_sym(cs, bw)
    fr = P.irsoft(cs)
    assert np.abs(P.isoft(cs).real - fr).max() < 1e-13
    assert rel_l2(P.rsoft(fr), cs) < 1e-13
    g = rng.standard_normal(P.grid_shape)
    assert rel_l2(P.rsoft(g), P.soft(g + 0j)) < 1e-13


def test_many_and_integrate():
    bw = 6
    rng = np.random.default_rng(3)
    P = O.OraclePlan(bw)
    cs = rng.standard_normal((3, P.nc)) + 1j * rng.standard_normal((3, P.nc))
    fs = P.isoft_many(cs)
    for i in range(3):
        np.testing.assert_array_equal(fs[i], P.isoft(cs[i]))
    np.testing.assert_allclose(P.soft_many(fs), cs, atol=1e-12)
    v = P.integrate(np.ones(P.grid_shape, np.complex128))
    assert abs(v - 8 * np.pi ** 2) < 1e-10


def test_fft2_vs_numpy():
    bw = 6
    rng = np.random.default_rng(4)
    P = O.OraclePlan(bw)
    X = rng.standard_normal(P.grid_shape) + 1j * rng.standard_normal(P.grid_shape)
    np.testing.assert_allclose(P.fft2(X, -1), np.fft.fft2(X, axes=(0, 1)), atol=1e-12)
    np.testing.assert_allclose(P.fft2(X, +1), np.fft.ifft2(X, axes=(0, 1)) * 144, atol=1e-12)


def test_rotate_vs_jacobi():
    bw = 7
    rng = np.random.default_rng(5)
    y = rng.standard_normal(bw * bw) + 1j * rng.standard_normal(bw * bw)
    for e in ([0.3, 1.1, 2.0], [5.0, 2.9, -1.0], [0.0, 0.0, 0.0]):
        r = O.rotate_ylm_cmplx(y, e)[0, 0]
        np.testing.assert_allclose(r, brute.rotate_ylm_direct(y, e, bw), atol=1e-12)
    # group law: rotating by (a,b,g) then its inverse (-g,-b,-a) is the identity
    e = np.array([0.7, 1.2, 2.5])
    r = O.rotate_ylm_cmplx(O.rotate_ylm_cmplx(y, e)[0, 0], -e[::-1])[0, 0]
    np.testing.assert_allclose(r, y, atol=1e-12)


def test_correlation_argmax_recovers_rotation():
    """tests/test_fortran_backend.py:695-722: argmax of |cc| is the grid rotation."""
    bw = 8
    P = O.OraclePlan(bw)
    rng = np.random.default_rng(6)
    f = rng.random(bw * bw) + 1j * rng.random(bw * bw)
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    for _ in range(5):
        idx = rng.integers(0, 2 * bw, 3)
        g = O.rotate_ylm_cmplx(f, [a[idx[0]], b[idx[1]], a[idx[2]]])[0, 0]
        cc = P.cross_correlation_ylm_cmplx(f, g)
        am = np.array(np.unravel_index(np.argmax(np.abs(cc.T)), cc.shape))
        assert list(am[[1, 0, 2]]) == list(idx)
    # correlation equals the isoft of its closed-form coefficients (soft.f08:2277, :2367)
    g = rng.random(bw * bw) + 1j * rng.random(bw * bw)
    assert rel_l2(P.cross_correlation_ylm_cmplx(f, g), P.isoft(O.corr_coeffs(bw, f, g))) < 1e-14
    # plant a mistake in the oracle's input: conj(g) instead of g must not pass
    assert rel_l2(P.cross_correlation_ylm_cmplx(f, np.conj(g)), P.isoft(O.corr_coeffs(bw, f, g))) > 1e-3


@pytest.mark.skipif(not RefLib.available(), reason="reference build oracle/_ref not present")
def test_live_reference_build_agrees_risbo():
    """The reduced Risbo step at arbitrary symmetric beta sets and the tables at more
    bandwidths, live against the reference build."""
    rng = np.random.default_rng(5)
    b = np.sort(rng.uniform(0.05, np.pi / 2 - 0.05, 5))
    b = np.concatenate([b, np.pi - b[::-1]])
    c, sn = np.cos(b / 2), np.sin(b / 2)
    d = dr = None
    for l in range(14):
        d = O.risbo_reduced_step(d, l, c, sn, True)
        dr = RefLib.risbo_reduced_step(dr, l, c, sn, True)
        np.testing.assert_array_equal(d, dr)
    for bw in (2, 3, 11, 16):
        np.testing.assert_array_equal(O.genwig_all_risbo(bw), RefLib.genwig_all_risbo(bw))
        np.testing.assert_array_equal(O.genwig_all_kostelec(bw), RefLib.genwig_all_kostelec(bw))
    for bw in (1, 2, 5, 9):
        np.testing.assert_allclose(O.full_big_d_risbo(bw, 1.9, 0.4, 5.1), RefLib.full_big_d_norm(bw, 1.9, 0.4, 5.1),
                                   rtol=0, atol=1e-15)


@pytest.mark.skipif(not RefLib.available(), reason="reference build oracle/_ref not present")
def test_live_reference_build_agrees():
    for bw in [3, 9, 17]:
        assert np.array_equal(O.coeff_degrees(bw), RefLib.coeff_degrees(bw))
        np.testing.assert_array_equal(O.seeds(bw), RefLib.seeds(bw))
        for m, n in [(0, 0), (-2, 1), (bw - 1, -(bw - 1)), (1, -bw + 1)]:
            for l in range(max(abs(m), abs(n)), bw):
                assert O.slice_table(bw)[m + bw - 1, n + bw - 1] + l - max(abs(m), abs(n)) == \
                    RefLib.coeff_location_mnl(m, n, l, bw)


def test_correlation_scale_at_true_rotation():
    """Scale and phase of the correlation, pinned by Parseval: for g = D(R0) f with R0 a
    grid rotation, C(R0) = <D(R0) f, g> = sum_lm |f_lm|^2 (real, positive).  A wrong
    normalisation 2 pi sqrt(2/(2l+1)) or sign (-1)^(m+n) in the coefficient product
    (soft.f08:2367) breaks the equality."""
    bw = 8
    P = O.OraclePlan(bw)
    rng = np.random.default_rng(61)
    f = rng.standard_normal(bw * bw) + 1j * rng.standard_normal(bw * bw)
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    for idx in ([3, 5, 11], [0, 0, 0], [15, 9, 2]):
        g = O.rotate_ylm_cmplx(f, [a[idx[0]], b[idx[1]], a[idx[2]]])[0, 0]
        cc = P.cross_correlation_ylm_cmplx(f, g)
        v = cc[idx[2], idx[0], idx[1]]
        assert abs(v - np.sum(np.abs(f) ** 2)) < 1e-11 * np.sum(np.abs(f) ** 2), (idx, v)


def _real_mlr(rng, bw, shape=()):
    """Random MLr coefficients of a real function (the m = 0 entries real)."""
    nr = bw * (bw + 1) // 2
    f = rng.standard_normal(shape + (nr,)) + 1j * rng.standard_normal(shape + (nr,))
    f[..., :bw] = f[..., :bw].real
    return f


def test_real_correlation_restatement_vs_complex_correlation():
    """The loop-by-loop restatement of the reference's half-complex real correlation
    (soft.f08:2806-2946, 3118-3131) equals the complex correlation of the same functions
    (LMc with f^l_{-m} = (-1)^m conj f^l_m), whose imaginary part vanishes; a wrong sign or
    beta reversal in any of the four half-complex branches breaks the equality."""
    for bw in (1, 2, 5, 8):
        rng = np.random.default_rng(70 + bw)
        f, g = _real_mlr(rng, bw), _real_mlr(rng, bw)
        P = O.OraclePlan(bw)
        cr = P.cross_correlation_ylm_real(f, g)
        cc = P.cross_correlation_ylm_cmplx(O.mlr_to_lmc(f, bw), O.mlr_to_lmc(g, bw))
        assert rel_l2(cr, cc.real) < 1e-12, bw
        assert np.abs(cc.imag).max() < 1e-12 * np.abs(cc).max(), bw


def test_real_correlation_peak_at_rotation():
    """tests/test_fortran_backend.py:723-750 on the oracle: the real correlation of f and
    its rotation by a grid rotation peaks at that rotation's grid index."""
    bw = 8
    P = O.OraclePlan(bw)
    rng = np.random.default_rng(77)
    f = _real_mlr(rng, bw)
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    for idx in ([3, 5, 11], [14, 2, 7]):
        gl = O.rotate_ylm_cmplx(O.mlr_to_lmc(f, bw), [a[idx[0]], b[idx[1]], a[idx[2]]])[0, 0]
        g = np.array([gl[l * (l + 1) + m] for m in range(bw) for l in range(m, bw)])
        cr = P.cross_correlation_ylm_real(f, g)
        ig, ia, ib = np.unravel_index(np.argmax(np.abs(cr)), cr.shape)
        assert [ia, ib, ig] == idx


def test_mlp_oracle_gradients_vs_finite_differences():
    """The alignment MLP's FP64 oracle gradients (oracle/mlp_oracle.py) against central
    differences of its own loss on a tiny network: a dropped chain-rule term or a wrong
    sign in the Gram-Schmidt backward would fail."""
    from oracle import mlp_oracle as M
    rng = np.random.default_rng(9)
    dims = (7, 5, 4, 6)
    n = 6
    sizes = sum(dims[i] * dims[i + 1] + dims[i + 1] for i in range(len(dims) - 1))
    p = rng.standard_normal(sizes) * 0.5
    X = rng.standard_normal((n, dims[0]))
    Q, _ = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    T = Q.reshape(n, 9)
    g = M.gradients(p, dims, X, T)
    eps = 1e-6
    fd = np.empty_like(p)
    for i in range(p.size):
        e = np.zeros_like(p)
        e[i] = eps
        fd[i] = (M.loss(p + e, dims, X, T) - M.loss(p - e, dims, X, T)) / (2 * eps)
    assert np.max(np.abs(g - fd)) < 1e-7 * max(1.0, np.max(np.abs(fd)))
