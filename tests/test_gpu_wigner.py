"""GPU parity of the Wigner-d utilities (pysofft.wigner namespace) and of the transforms
that read the plan's d table (precompute_wigners, Risbo handles), against the oracle and
the reference build's golden vectors.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def test_genwig_all_tables_vs_golden_and_oracle(golden):
    """so3ft_genwig_all: the d table by the reduced Risbo recurrence (wigner.f90:784-822) and
    by the Kostelec recurrence (wigner.f90:423-465), [rows][2bw] on the device."""
    from pysofft_amd import wigner as Wg
    # entries reach sqrt((2l+1)/2) ~ 3: 1e-14 is a few ulps (the device contracts products
    # into fused multiply-adds, the reference build does not)
    np.testing.assert_allclose(Wg.genwig_all_risbo(9).T, golden["genwig_all_risbo_9"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(Wg.genwig_all_kostelec(9).T, golden["genwig_all_kostelec_9"], rtol=0, atol=1e-14)
    np.testing.assert_allclose(Wg.genwig_all_risbo(5, normalized=False).T, golden["genwig_all_risbo_raw_5"],
                               rtol=0, atol=1e-14)
    # at bw = 33 entries reach ~6 after 33 recurrence steps: 1e-13 is a few ulps of them
    for bw in (1, 2, 33):
        np.testing.assert_allclose(Wg.genwig_all_risbo(bw).T, O.genwig_all_risbo(bw), rtol=0, atol=1e-13)
        np.testing.assert_allclose(Wg.genwig_all_kostelec(bw).T, O.genwig_all_kostelec(bw), rtol=0, atol=1e-13)


def test_full_big_d_risbo_vs_golden_and_oracle(golden):
    from pysofft_amd import wigner as Wg
    al, be, ga = golden["full_big_d_angles"]
    np.testing.assert_allclose(Wg.full_big_d_risbo(6, al, be, ga), golden["full_big_d_6"], rtol=0, atol=1e-15)
    np.testing.assert_allclose(Wg.full_big_d_risbo(4, al, be, ga, normalized=False), golden["full_big_d_raw_4"],
                               rtol=0, atol=1e-15)
    x = Wg.full_big_d_risbo(20, 4.1, 2.2, 0.7)
    assert rel_l2(x, O.full_big_d_risbo(20, 4.1, 2.2, 0.7)) < 1e-14


def test_wigner_mn_recurrence_risbo_vs_golden_and_oracle(golden):
    """Full-matrix Risbo steps (wigner.f90:529-585): eleven steps at the golden beta equal the
    reference's d^11; a batch of betas equals the oracle's step, beta by beta."""
    from pysofft_amd import wigner as Wg
    cb, sb = golden["risbo_cs"]
    sq = np.sqrt(np.arange(25))
    d = None
    for l in range(12):
        d = Wg.wigner_mn_recurrence_risbo(d, l, np.array([cb]), np.array([sb]), sq)
    np.testing.assert_allclose(d[0], golden["risbo_l11"], rtol=0, atol=1e-14)
    rng = np.random.default_rng(3)
    b = rng.uniform(0, np.pi, 7)
    c, s = np.cos(b / 2), np.sin(-b / 2)
    sq = np.sqrt(np.arange(41))
    d = None
    ref = [None] * 7
    for l in range(20):
        d = Wg.wigner_mn_recurrence_risbo(d, l, c, s, sq)
        for k in range(7):
            ref[k] = O.risbo_step(ref[k], l, c[k], s[k], sq)
    for k in range(7):
        np.testing.assert_allclose(d[k], ref[k], rtol=0, atol=1e-13)


def test_wigner_recurrence_risbo_reduced_vs_oracle():
    """Reduced Risbo steps (wigner.f90:626-748) at a symmetric beta set, normalised and not,
    against the oracle's step; wigner_dl_risbo_reduced against the oracle table."""
    from pysofft_amd import wigner as Wg
    rng = np.random.default_rng(4)
    h = np.sort(rng.uniform(0.02, np.pi / 2 - 0.02, 6))
    b = np.concatenate([h, np.pi - h[::-1]])
    c, s = np.cos(b / 2), np.sin(b / 2)
    sq = np.sqrt(np.arange(80))
    for norm in (True, False):
        d = dr = None
        for l in range(30):
            d = Wg.wigner_recurrence_risbo_reduced(d, l, c, s, sq, norm)
            dr = O.risbo_reduced_step(dr, l, c, s, norm)
            assert d.shape == (len(b), (l + 1) * (l + 2) // 2)
            np.testing.assert_allclose(d.T, dr, rtol=0, atol=1e-13 * max(1.0, np.abs(dr).max()))
    bw = 10
    t = O.genwig_all_risbo(bw)
    bs = O.beta_samples(2 * bw)
    dl = Wg.wigner_dl_risbo_reduced(bw - 1, bs)
    # the degree-(bw-1) entries of the table are the last row of each pair's block
    r = 0
    for p in range(bw):
        for q in range(p, bw):
            r += bw - q
            np.testing.assert_allclose(dl[:, p * (2 * bw - p + 1) // 2 + q - p], t[r - 1], rtol=0, atol=1e-13)


@pytest.mark.parametrize("bw", [1, 2, 5, 16, 33])
def test_precomputed_kostelec_table_transforms_vs_oracle(bw):
    """precompute_wigners=True on a Kostelec handle: the FP64 transforms read the d table
    (genwig_all_kostelec, soft.f08:194-219) and equal the oracle; results are the same for
    both engine settings and for batches."""
    from pysofft_amd import Soft
    rng = np.random.default_rng(60 + bw)
    s = Soft(bw, precompute_wigners=True)
    assert s.wigners_are_precomputed and s.coeff_order == 'mnl'
    P = O.OraclePlan(bw)
    c = rng.standard_normal((3, P.nc)) + 1j * rng.standard_normal((3, P.nc))
    g = rng.standard_normal((3,) + P.grid_shape) + 1j * rng.standard_normal((3,) + P.grid_shape)
    for eng in ("valu", "mfma"):
        s.dwt_engine = eng
        assert rel_l2(np.asarray(s.isoft_many(c)), P.isoft_many(c)) < 1e-12
        assert rel_l2(np.asarray(s.soft_many(g)), P.soft_many(g)) < 1e-12
        assert rel_l2(np.asarray(s.rsoft_many(g.real)), P.rsoft_many(g.real)) < 1e-12
    cr = c[0].copy()
    O.enforce_real_sym(cr, bw)
    assert rel_l2(s.irsoft(cr), P.irsoft(cr)) < 1e-12
    fl = rng.standard_normal(bw * bw) + 1j * rng.standard_normal(bw * bw)
    gl = rng.standard_normal(bw * bw) + 1j * rng.standard_normal(bw * bw)
    assert rel_l2(s.cross_correlation_ylm_cmplx(fl, gl), P.cross_correlation_ylm_cmplx(fl, gl)) < 1e-12


@pytest.mark.parametrize("bw", [3, 16, 32])
def test_risbo_table_transforms_vs_oracle(bw):
    """Risbo handles (lmn order, d from the Risbo table) and Risbo handles with precomputed
    Wigners (mnl order, soft.f08:194-219) against the oracle."""
    from pysofft_amd import Soft, utils
    rng = np.random.default_rng(70 + bw)
    P = O.OraclePlan(bw)
    perm = utils.lmn_to_mnl(bw)
    c = rng.standard_normal(P.nc) + 1j * rng.standard_normal(P.nc)
    g = rng.standard_normal(P.grid_shape) + 1j * rng.standard_normal(P.grid_shape)
    s = Soft(bw, recurrence_type=1)
    cm = np.empty_like(c)
    cm[perm] = c
    assert rel_l2(s.isoft(c), P.isoft(cm)) < 1e-12
    assert rel_l2(np.asarray(s.soft(g)), P.soft(g)[perm]) < 1e-12
    sp = Soft(bw, recurrence_type=1, precompute_wigners=True)
    assert sp.coeff_order == 'mnl'
    assert rel_l2(sp.isoft(c), P.isoft(c)) < 1e-12
    assert rel_l2(np.asarray(sp.soft(g)), P.soft(g)) < 1e-12
