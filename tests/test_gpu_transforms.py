"""GPU parity of the HIP transform path against the CPU oracle (FP64).

Bar: relative L2 <= 1e-12 for every complex/real FP64 transform against the oracle
on the same seeded inputs; round trips at sizes the oracle cannot finish quickly
are checked through size-independent properties.
"""
import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _soft(bw, **kw):
    from pysofft_amd import Soft
    return Soft(bw, **kw)


def _rand_c(rng, n):
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


@pytest.mark.parametrize("bw", [1, 2, 3, 4, 5, 7, 8, 13, 16, 31, 32])
def test_isoft_soft_vs_oracle(bw):
    rng = np.random.default_rng(1000 + bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    f = s.isoft(c)
    fo = P.isoft(c)
    assert rel_l2(f, fo) < 1e-12, bw
    g = _rand_c(rng, P.grid_shape)
    assert rel_l2(np.asarray(s.soft(g)), P.soft(g)) < 1e-12, bw


def test_workload_bw32_round_trip():
    """bw=32 complex round trip on iid N(0,1) coefficients: isoft then soft, both vs oracle."""
    bw = 32
    rng = np.random.default_rng(32)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    f = s.isoft(c)
    c2 = np.asarray(s.soft(f))
    assert rel_l2(f, P.isoft(c)) <= 1e-12
    assert rel_l2(c2, P.soft(P.isoft(c))) <= 1e-12
    assert rel_l2(c2, c) <= 1e-12


@pytest.mark.parametrize("bw", [48, 64])
def test_isoft_soft_vs_oracle_medium(bw):
    rng = np.random.default_rng(bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    f = s.isoft(c)
    assert rel_l2(f, P.isoft(c)) < 1e-12
    assert rel_l2(np.asarray(s.soft(f)), c) < 1e-12


@pytest.mark.parametrize("bw", [96, 128])
def test_large_bw_round_trip_and_samples(bw):
    """FP64 round trip at bw = 3 * 2^5 and 2^7 (FFT lengths 192 and 256) and grid samples of
    the l <= 6 part against a direct sum."""
    rng = np.random.default_rng(bw)
    s = _soft(bw)
    nc = s.n_coeffs
    c = torch.from_numpy(_rand_c(rng, nc)).cuda()
    f = s.isoft(c)
    c2 = s.soft(f)
    assert rel_l2(c2.cpu().numpy(), c.cpu().numpy()) < 1e-12
    # grid values of the l <= 6 part against a direct sum with Jacobi-formula d (oracle/brute.py)
    deg = s.coeff_indices
    cn = c.cpu().numpy()
    fn = f.cpu().numpy()
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    N = 2 * bw
    pts = [(0, 0, 0), (5, 17, 100), (N - 1, bw, 3), (77, N - 56, N - 1)]
    from oracle import brute
    sel = [i for i in range(len(deg)) if deg[i][0] <= 6]
    cl = np.zeros(nc, np.complex128)
    cl[sel] = cn[sel]
    fl = s.isoft(cl)
    for (ig, ia, ib) in pts:
        v = 0j
        for i in sel:
            l, m, n = deg[i]
            v += cl[i] * np.exp(-1j * m * a[ia]) * brute.wigner_d_norm(l, m, n, b[ib:ib + 1])[0] * \
                np.exp(-1j * n * a[ig]) / (2 * np.pi)
        assert abs(fl[ig, ia, ib] - v) < 1e-12 * max(1, abs(v))
    assert np.isfinite(fn).all()


@pytest.mark.parametrize("bw", [1, 2, 3, 6, 9, 16, 32])
def test_real_transforms_vs_oracle(bw):
    rng = np.random.default_rng(2000 + bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    O.enforce_real_sym(c, bw)
    fr = s.irsoft(c)
    assert fr.dtype == np.float64
    assert rel_l2(fr, P.irsoft(c)) < 1e-12
    g = rng.standard_normal(P.grid_shape)
    cr = np.asarray(s.rsoft(g))
    assert rel_l2(cr, P.rsoft(g)) < 1e-12
    assert rel_l2(np.asarray(s.rsoft(fr)), c) < 1e-12


def test_many_equals_singles():
    bw = 9
    rng = np.random.default_rng(9)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    cs = _rand_c(rng, (5, P.nc))
    fs = s.isoft_many(cs)
    assert fs.shape == (5,) + P.grid_shape
    for i in range(5):
        assert rel_l2(fs[i], P.isoft(cs[i])) < 1e-12
    c2 = np.asarray(s.soft_many(fs))
    assert rel_l2(c2, cs) < 1e-12
    csr = cs.copy()
    for i in range(5):
        O.enforce_real_sym(csr[i], bw)
    fr = s.irsoft_many(csr)
    assert rel_l2(fr, P.irsoft_many(csr)) < 1e-12
    assert rel_l2(np.asarray(s.rsoft_many(fr)), csr) < 1e-12


def test_torch_inputs_and_out_semantics():
    bw = 8
    rng = np.random.default_rng(8)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    ct = torch.from_numpy(c).cuda()
    f = s.isoft(ct)
    assert isinstance(f, torch.Tensor) and f.is_cuda and f.shape == P.grid_shape
    out = s.get_so3func()
    r = s.isoft(c, out=out)
    assert r is out
    assert rel_l2(out, f.cpu().numpy()) < 1e-15
    co = s.soft(out)
    assert co.lmn[3, -2, 1].shape == (1,)
    assert abs(co.lmn[3, -2, 1][0] - c[s.coeff_indices.tolist().index([3, -2, 1])]) < 1e-12


@pytest.mark.parametrize("bw,lmax", [(8, 5), (9, 0), (16, 11)])
def test_lmax_below_bw(bw, lmax):
    rng = np.random.default_rng(lmax)
    s = _soft(bw, lmax=lmax)
    P = O.OraclePlan(bw, lmax)
    c = _rand_c(rng, P.nc)
    assert rel_l2(s.isoft(c), P.isoft(c)) < 1e-12
    g = _rand_c(rng, P.grid_shape)
    co = np.asarray(s.soft(g))
    deg = s.coeff_indices
    mask = np.maximum(np.abs(deg[:, 1]), np.abs(deg[:, 2])) <= lmax
    assert rel_l2(co[mask], P.soft(g)[mask]) < 1e-12


@pytest.mark.parametrize("bw", [4, 11])
def test_stage_entry_points(bw):
    rng = np.random.default_rng(bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, P.nc)
    F = s._inverse_wigner_trf_cmplx(c)
    assert rel_l2(F, P.inverse_dwt(c)) < 1e-12
    G = _rand_c(rng, P.grid_shape)
    assert rel_l2(np.asarray(s._forward_wigner_trf_cmplx(G)), P.forward_dwt(G)) < 1e-12
    # fft = FFTW_FORWARD 2-D over (alpha, gamma) / (2bw)  (soft.f08:3256-3262)
    N = 2 * bw
    assert rel_l2(s._fft(G), np.fft.fft2(G, axes=(0, 1)) / N) < 1e-13
    assert rel_l2(s._ifft(G), np.fft.ifft2(G, axes=(0, 1)) * N) < 1e-13


def test_integrate_over_so3():
    for bw in [1, 5, 64]:
        s = _soft(bw)
        one = np.ones(s.grid_shape)
        assert abs(s.integrate_over_so3(one) - 8 * np.pi ** 2) < 1e-10
        assert abs(s.integrate_over_so3(one + 0j) - 8 * np.pi ** 2) < 1e-10


@pytest.mark.parametrize("bw", [8, 16])
def test_cross_correlation_vs_oracle_and_argmax(bw):
    rng = np.random.default_rng(bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    f = rng.random(bw * bw) + 1j * rng.random(bw * bw)
    g = rng.random(bw * bw) + 1j * rng.random(bw * bw)
    assert rel_l2(s.cross_correlation_ylm_cmplx(f, g), P.cross_correlation_ylm_cmplx(f, g)) < 1e-12
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    for _ in range(4):
        idx = rng.integers(0, 2 * bw, 3)
        gr = s.rotate_ylm_cmplx(f, (a[idx[0]], b[idx[1]], a[idx[2]]))
        cc = s.cross_correlation_ylm_cmplx(f, gr)
        am = np.array(np.unravel_index(np.argmax(np.abs(cc.T)), cc.shape))
        assert list(am[[1, 0, 2]]) == list(idx)


def test_cross_correlation_3d_vs_oracle():
    bw, nr = 8, 4
    rng = np.random.default_rng(77)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    f = rng.random((nr, bw * bw)) + 1j * rng.random((nr, bw * bw))
    g = rng.random((nr, bw * bw)) + 1j * rng.random((nr, bw * bw))
    rad = np.arange(nr).astype(float) + 0.5
    got = s.cross_correlation_ylm_cmplx_3d(f, g, radial_sampling_points=rad, radial_limits=[1, nr])
    assert rel_l2(got, P.cross_correlation_ylm_cmplx_3d(f, g, rad, (1, nr))) < 1e-12


def _real_mlr(rng, bw, shape=()):
    nr = bw * (bw + 1) // 2
    f = rng.standard_normal(shape + (nr,)) + 1j * rng.standard_normal(shape + (nr,))
    f[..., :bw] = f[..., :bw].real
    return f


@pytest.mark.parametrize("engine", ["valu", "mfma"])
@pytest.mark.parametrize("bw", [1, 2, 5, 8, 13])
def test_real_correlation_vs_oracle(bw, engine):
    """cross_correlation_ylm_real (MLr inputs, float64 grid) and its DWT stage against
    the oracle's loop-by-loop restatement of soft.f08:2806-2946, 3118-3131."""
    rng = np.random.default_rng(100 + bw)
    s = _soft(bw)
    s.dwt_engine = engine
    P = O.OraclePlan(bw)
    f, g = _real_mlr(rng, bw), _real_mlr(rng, bw)
    got = s.cross_correlation_ylm_real(f, g)
    assert got.dtype == np.float64 and got.shape == tuple(P.grid_shape)
    assert rel_l2(got, P.cross_correlation_ylm_real(f, g)) < 1e-12
    st = s._inverse_wigner_trf_corr_real(f, g)
    assert rel_l2(st, P.cross_correlation_ylm_real(f, g, stage=True)) < 1e-12
    # the complex stage against the oracle's inverse DWT of its correlation coefficients
    fc, gc = _rand_c(rng, bw * bw), _rand_c(rng, bw * bw)
    st = s._inverse_wigner_trf_corr_cmplx(fc, gc)
    assert rel_l2(st, P.inverse_dwt(O.corr_coeffs(bw, fc, gc))) < 1e-12


def test_real_correlation_3d_vs_oracle():
    bw, nr = 8, 5
    rng = np.random.default_rng(78)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    f, g = _real_mlr(rng, bw, (nr,)), _real_mlr(rng, bw, (nr,))
    rad = np.arange(nr).astype(float) + 0.25
    got = s.cross_correlation_ylm_real_3d(f, g, radial_sampling_points=rad, radial_limits=[1, nr])
    assert rel_l2(got, P.cross_correlation_ylm_real_3d(f, g, rad, (1, nr))) < 1e-12


def test_real_correlation_peak_bw32():
    """tests/test_fortran_backend.py:723-750: real f (rng.random MLr coefficients) and
    its rotation by grid rotations (rotate_ylm_real); the peak of the real correlation is
    the rotation's grid index."""
    from pysofft_amd import Soft
    bw = 32
    rng = np.random.default_rng(32)
    nr = bw * (bw + 1) // 2
    f = rng.random(nr).astype(complex)
    s = _soft(bw)
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    for _ in range(5):
        idx = rng.integers(0, 2 * bw, 3)
        g = Soft.rotate_ylm_real(f[None], (a[idx[0]], b[idx[1]], a[idx[2]]))
        cc = s.cross_correlation_ylm_real(f, np.asarray(g).reshape(-1)[:nr])
        ig, ia, ib = np.unravel_index(np.argmax(np.abs(cc)), cc.shape)
        assert [ia, ib, ig] == list(idx)


def test_rotate_ylm_vs_oracle():
    bw = 12
    rng = np.random.default_rng(12)
    from pysofft_amd import Soft
    y = _rand_c(rng, (3, bw * bw))
    e = np.stack([rng.uniform(0, 2 * np.pi, 4), rng.uniform(0, np.pi, 4), rng.uniform(0, 2 * np.pi, 4)], 1)
    got = Soft.rotate_ylm_cmplx(y, e)
    ref = O.rotate_ylm_cmplx(y, e)
    assert got.shape == (4, 3, bw * bw)
    assert rel_l2(got, ref) < 1e-12


def test_enforce_real_symmetry_on_device(golden):
    bw = 6
    s = _soft(bw)
    c = torch.from_numpy(golden["sym_in_6"].copy()).cuda()
    s.enforce_real_symmetry(c)
    np.testing.assert_array_equal(c.cpu().numpy(), golden["sym_out_6"])


def test_invalid_inputs_raise_not_abort():
    s = _soft(4)
    with pytest.raises(ValueError):
        s.isoft(np.zeros(10, np.complex128))
    from pysofft_amd._lib import So3ftError
    with pytest.raises(So3ftError):
        _soft(513)


@pytest.mark.parametrize("bw", [1, 2, 5, 16, 33])
@pytest.mark.parametrize("n", [1, 3, 6, 17])
def test_dwt_engines_agree_with_oracle(bw, n):
    """Both FP64 DWT engines (vector FMA and MFMA) against the oracle for batch sizes that
    exercise every transform grouping of the batched kernels (1, 3, 6, 17 transforms)."""
    rng = np.random.default_rng(7 * bw + n)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    c = _rand_c(rng, (n, P.nc))
    g = _rand_c(rng, (n,) + tuple(P.grid_shape))
    fo, co = P.isoft_many(c), P.soft_many(g)
    for eng in ("valu", "mfma"):
        s.dwt_engine = eng
        assert s.dwt_engine == eng
        assert rel_l2(np.asarray(s.isoft_many(c)), fo) < 1e-12, eng
        assert rel_l2(np.asarray(s.soft_many(g)), co) < 1e-12, eng
        rr = np.asarray(s.rsoft_many(g.real))
        assert rel_l2(rr, P.rsoft_many(g.real)) < 1e-12, eng
        cr = np.stack([O.enforce_real_sym(ci.copy(), bw) for ci in c])
        assert rel_l2(np.asarray(s.irsoft_many(cr)), P.irsoft_many(cr)) < 1e-12, eng


@pytest.mark.parametrize("bw", [2, 3, 8, 12, 16, 24, 32, 48, 64])
def test_mixed_c64_vs_oracle(bw):
    """complex64 transforms (FP32 MFMA contraction, FP64 recurrence and FFTs) against the
    FP64 oracle on the same complex64-rounded inputs: per-transform relative L2 <= 1e-7
    for isoft, soft and the round trip.  Five transforms: at bw = 64 the forward runs four
    per wave and a one-transform tail."""
    rng = np.random.default_rng(500 + bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    n = 5
    c = _rand_c(rng, (n, P.nc)).astype(np.complex64)
    f = np.asarray(s.isoft_many_c64(c))
    assert f.dtype == np.complex64
    fo = P.isoft_many(c.astype(np.complex128))
    g = _rand_c(rng, (n,) + tuple(P.grid_shape)).astype(np.complex64)
    co = P.soft_many(g.astype(np.complex128))
    cg = np.asarray(s.soft_many_c64(g))
    c2 = np.asarray(s.soft_many_c64(f))
    for i in range(n):
        assert rel_l2(f[i], fo[i]) <= 1e-7, (bw, i, rel_l2(f[i], fo[i]))
        assert rel_l2(cg[i], co[i]) <= 1e-7, (bw, i, rel_l2(cg[i], co[i]))
        assert rel_l2(c2[i], c[i]) <= 2e-7, (bw, i, rel_l2(c2[i], c[i]))


@pytest.mark.parametrize("bw,n", [(96, 5), (128, 5), (128, 2), (256, 1)])
def test_mixed_c64_large_bw_vs_fp64(bw, n):
    """complex64 transforms at bandwidths beyond the oracle's reach (FFT lengths 192, 256,
    512) against the FP64 path, itself pinned to the oracle at smaller bandwidths and to
    direct sums at bw = 96 and 128: relative L2 <= 1e-7 per transform for isoft and soft.
    n = 5 runs the forward with four transforms per wave (bw >= 64) and its one-transform
    tail; n = 2 and 1 the two- and one-transform waves."""
    N = 2 * bw
    s = _soft(bw)

    def rel_l2(a, b):   # on the device: the grids hold up to 2^27 points
        a = a.to(torch.complex128)
        return ((a - b).abs().pow(2).sum() / b.abs().pow(2).sum()).sqrt().item()

    g = torch.Generator(device="cuda").manual_seed(bw)
    c32 = torch.view_as_complex(torch.randn(n, s.n_coeffs, 2, generator=g, device="cuda")).contiguous()
    f32 = s.isoft_many_c64(c32)
    f64 = s.isoft_many(c32.to(torch.complex128))
    for i in range(n):
        e = rel_l2(f32[i], f64[i])
        assert e <= 1e-7, (bw, i, e)
    del f64
    r32 = s.soft_many_c64(f32)
    r64 = s.soft_many(f32.to(torch.complex128))
    for i in range(n):
        e = rel_l2(r32[i], r64[i])
        assert e <= 1e-7, (bw, i, e)
    assert f32.shape == (n, N, N, N)


@pytest.mark.parametrize("bw", [5, 7, 20, 40])
def test_mixed_c64_any_bw_vs_oracle(bw):
    """complex64 transforms at bandwidths that are not 2^a or 3 * 2^a (FFT lengths 10, 14,
    40, 80: the Stockham pass with complex64 storage) against the FP64 oracle:
    per-transform relative L2 <= 1e-7."""
    rng = np.random.default_rng(800 + bw)
    s = _soft(bw)
    P = O.OraclePlan(bw)
    n = 3
    c = _rand_c(rng, (n, P.nc)).astype(np.complex64)
    f = np.asarray(s.isoft_many_c64(c))
    fo = P.isoft_many(c.astype(np.complex128))
    g = _rand_c(rng, (n,) + tuple(P.grid_shape)).astype(np.complex64)
    co = P.soft_many(g.astype(np.complex128))
    cg = np.asarray(s.soft_many_c64(g))
    for i in range(n):
        assert rel_l2(f[i], fo[i]) <= 1e-7, (bw, i, rel_l2(f[i], fo[i]))
        assert rel_l2(cg[i], co[i]) <= 1e-7, (bw, i, rel_l2(cg[i], co[i]))


@pytest.mark.parametrize("engine,n", [("valu", 5), ("valu", 2), ("mfma", 5)])
def test_cross_correlation_many_vs_oracle(engine, n):
    """Batched correlation vs the oracle: n = 5 runs four transforms per wave plus a tail of
    one, n = 2 two per wave (the FP64 inverse forms the coefficient products as it loads)."""
    bw = 10
    rng = np.random.default_rng(31)
    s = _soft(bw)
    s.dwt_engine = engine
    P = O.OraclePlan(bw)
    f = _rand_c(rng, (n, bw * bw))
    g = _rand_c(rng, (n, bw * bw))
    got = np.asarray(s.cross_correlation_ylm_cmplx_many(f, g))
    assert got.shape == (n,) + tuple(P.grid_shape)
    for i in range(n):
        assert rel_l2(got[i], P.cross_correlation_ylm_cmplx(f[i], g[i])) < 1e-12, i


@pytest.mark.parametrize("bw", [1, 2, 3])
@pytest.mark.parametrize("engine", ["valu", "mfma"])
def test_cross_correlation_smallest_bandwidths_vs_oracle(bw, engine):
    """Correlation at the smallest bandwidths (one to three degrees, chains of one degree) vs
    the oracle, single and batched (three pairs: one wave's tail)."""
    rng = np.random.default_rng(90 + bw)
    s = _soft(bw)
    s.dwt_engine = engine
    P = O.OraclePlan(bw)
    f = _rand_c(rng, (3, bw * bw))
    g = _rand_c(rng, (3, bw * bw))
    assert rel_l2(s.cross_correlation_ylm_cmplx(f[0], g[0]), P.cross_correlation_ylm_cmplx(f[0], g[0])) < 1e-12
    got = np.asarray(s.cross_correlation_ylm_cmplx_many(f, g))
    for i in range(3):
        assert rel_l2(got[i], P.cross_correlation_ylm_cmplx(f[i], g[i])) < 1e-12, i


def test_correlation_peaks_find_grid_rotations():
    """g = exact Wigner-D rotation of f by a grid rotation: the fused correlation + peak
    search returns that rotation's grid indices (the reference's own check,
    tests/test_fortran_backend.py:695-722), and max|C| equals the grid maximum."""
    bw, n = 16, 12
    rng = np.random.default_rng(16)
    from pysofft_amd import Soft
    s = _soft(bw)
    a = np.arange(2 * bw) * np.pi / bw
    b = O.beta_samples(2 * bw)
    f = rng.random((n, bw * bw)) + 1j * rng.random((n, bw * bw))
    idx = rng.integers(0, 2 * bw, (n, 3))
    e = np.stack([a[idx[:, 0]], b[idx[:, 1]], a[idx[:, 2]]], 1)
    g = Soft.rotate_ylm_cmplx(f, e)           # (n rotations, n sets, bw^2)
    g = np.stack([g[i, i] for i in range(n)])
    peak, val = s.correlation_peaks(f, g)
    np.testing.assert_array_equal(peak, idx)
    cc = np.asarray(s.cross_correlation_ylm_cmplx_many(f, g))
    np.testing.assert_allclose(val, np.abs(cc).reshape(n, -1).max(1), rtol=1e-14)


@pytest.mark.parametrize("bw", [5, 6, 24, 64, 96])
def test_correlation_peaks_equal_full_grid_argmax(bw):
    """The peak search (the last FFT pass fused with the search where that pass has a
    one-tile shape: 2bw = 12, 48, 128, 192 here; the stored grid and a separate search at
    2bw = 10) returns, per pair, the first C-order index of max |C| over the full
    correlation grid and that maximum, for random (not rotated) pairs."""
    rng = np.random.default_rng(900 + bw)
    s = _soft(bw)
    n = 3 if bw < 96 else 2
    f = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    g = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    ft, gt = torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda()
    peak, val = s.correlation_peaks(ft, gt)
    cc = s.cross_correlation_ylm_cmplx_many(ft, gt)
    m = cc.abs().reshape(n, -1)
    flat = m.argmax(1)
    N = 2 * bw
    ig, rem = flat // (N * N), flat % (N * N)
    ref = torch.stack([rem // N, rem % N, ig], 1)
    assert torch.equal(peak, ref)
    torch.testing.assert_close(val, m.max(1).values, rtol=1e-14, atol=0)


@pytest.mark.parametrize("bw,lmax,rt", [(24, 15, 0), (64, 40, 0), (16, 9, 1)])
def test_correlation_peaks_lmax_below_bw(bw, lmax, rt):
    """The fused peak search on handles that drop degrees (and orders) above lmax -- the
    last pass then reads only the rows of orders |m| <= lmax -- returns the argmax of the
    same handle's full correlation grids and that maximum (Kostelec and Risbo handles)."""
    rng = np.random.default_rng(700 + bw)
    s = _soft(bw, lmax=lmax, recurrence_type=rt)
    n = 3
    f = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    g = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    ft, gt = torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda()
    peak, val = s.correlation_peaks(ft, gt)
    cc = s.cross_correlation_ylm_cmplx_many(ft, gt)
    m = cc.abs().reshape(n, -1)
    flat = m.argmax(1)
    N = 2 * bw
    ig, rem = flat // (N * N), flat % (N * N)
    ref = torch.stack([rem // N, rem % N, ig], 1)
    assert torch.equal(peak, ref)
    torch.testing.assert_close(val, m.max(1).values, rtol=1e-14, atol=0)


def test_workload_bw256_real_round_trip():
    """Real FP64 transform at bw=256: Hermitian-symmetric coefficients (iid N(0,1) re/im
    scaled by (2l+1)^-1/2), irsoft onto the 512^3 float64 grid, rsoft back: relative L2
    <= 1e-11 for the round trip; grid samples of the l <= 3 part against direct sums with
    the Jacobi-formula d (oracle/brute.py)."""
    from pysofft_amd import Soft
    from oracle import brute
    bw = 256
    s = Soft(bw)
    rng = np.random.default_rng(256)
    nc = s.n_coeffs
    deg = O.coeff_degrees(bw)
    ls = deg[:, 0]
    c = (rng.standard_normal(nc) + 1j * rng.standard_normal(nc)) / np.sqrt(2 * ls + 1)
    O.enforce_real_sym(c, bw)
    f = s.irsoft(torch.from_numpy(c).cuda())
    assert f.dtype == torch.float64 and tuple(f.shape) == (512, 512, 512)
    c2 = s.rsoft(f)
    assert rel_l2(c2.cpu().numpy(), c) <= 1e-11
    del f, c2
    sel = np.nonzero(ls <= 3)[0]
    lo = np.zeros(nc, np.complex128)
    lo[sel] = c[sel]
    fl = s.irsoft(torch.from_numpy(lo).cuda())
    a, b, _ = brute.grid_angles(bw)
    for ig, ia, ib in [(0, 0, 0), (17, 200, 311), (511, 255, 3), (128, 384, 77)]:
        v = 0j
        for i in sel:
            l, m, n = deg[i]
            v += lo[i] * np.exp(-1j * m * a[ia]) * brute.wigner_d_norm(l, m, n, b[ib:ib + 1])[0] * \
                np.exp(-1j * n * a[ig]) / (2 * np.pi)
        assert abs(v.imag) < 1e-12
        got = fl[ig, ia, ib].item()
        assert abs(got - v.real) <= 1e-12 * max(1.0, abs(v)), (ig, ia, ib, got, v)


@pytest.mark.parametrize("bw", [1, 2, 5, 8, 16])
def test_risbo_lmn_layout_transforms_vs_oracle(bw):
    """recurrence_type = Risbo without precomputed Wigners: coefficients in lmn order
    (pysofft/soft.py:98-106, utils.f90:228-355) for every transform and stage."""
    from pysofft_amd import Soft, utils
    rng = np.random.default_rng(300 + bw)
    s = Soft(bw, recurrence_type=1)
    assert s.coeff_order == 'lmn' and s.recurrence_type == 1
    assert Soft(bw, recurrence_type=1, precompute_wigners=True).coeff_order == 'mnl'
    P = O.OraclePlan(bw)
    perm = utils.lmn_to_mnl(bw)
    c = _rand_c(rng, utils.total_num_coeffs(bw))
    cm = np.empty_like(c)
    cm[perm] = c                      # the same coefficients in mnl order
    f = s.isoft(c)
    fo = P.isoft(cm)
    assert rel_l2(f, fo) < 1e-12
    assert rel_l2(np.asarray(s.soft(fo)), P.soft(fo)[perm]) < 1e-12
    fd = s._inverse_wigner_trf_cmplx(c)
    assert rel_l2(fd, P.inverse_dwt(cm)) < 1e-12
    assert rel_l2(np.asarray(s._forward_wigner_trf_cmplx(fd)), P.forward_dwt(fd)[perm]) < 1e-12
    # many, real and mixed-precision entry points
    cs = _rand_c(rng, (3, utils.total_num_coeffs(bw)))
    fs = np.asarray(s.isoft_many(cs))
    for i in range(3):
        ci = np.empty_like(cs[i])
        ci[perm] = cs[i]
        assert rel_l2(fs[i], P.isoft(ci)) < 1e-12
    cr = c.copy()
    s.enforce_real_symmetry(cr)
    cmr = np.empty_like(cr)
    cmr[perm] = cr
    assert rel_l2(s.irsoft(cr), P.irsoft(cmr)) < 1e-12
    assert rel_l2(np.asarray(s.rsoft(P.irsoft(cmr))), cr) < 1e-11
    if bw in (2, 8, 16):  # complex64 transforms need bw = 2^a or 3*2^a
        import torch
        c32 = torch.from_numpy(cs.astype(np.complex64)).cuda()
        f32 = s.isoft_many_c64(c32)
        assert rel_l2(f32.cpu().numpy(), fs) < 1e-6
        assert rel_l2(s.soft_many_c64(f32).cpu().numpy(), c32.cpu().numpy()) < 1e-6


def test_enforce_real_sym_lmn_device_vs_host():
    import torch
    from pysofft_amd import Soft, utils
    bw = 9
    rng = np.random.default_rng(9)
    c = _rand_c(rng, (4, utils.total_num_coeffs(bw)))
    ref = utils.enforce_real_sym_lmn_many(c.copy(), bw)
    t = torch.from_numpy(c).cuda()
    Soft(bw, recurrence_type=1).enforce_real_symmetry(t)
    np.testing.assert_array_equal(t.cpu().numpy(), ref)


@pytest.mark.parametrize("bw", [1, 3, 8])
def test_rfft_irfft_vs_numpy(bw):
    """py_rfft / py_irfft (soft.f08:3270-3283): r2c over (alpha, gamma) with the alpha axis
    halved, c2r back (FFTW_BACKWARD, reading the stored half as FFTW does), both /(2bw);
    numpy's rfftn / irfftn over axes (gamma, alpha) are the library routines they equal."""
    N = 2 * bw
    rng = np.random.default_rng(bw)
    s = _soft(bw)
    f = rng.standard_normal((N, N, N))
    got = s._rfft(f)
    assert got.shape == (N, bw + 1, N)
    assert rel_l2(got, np.fft.rfftn(f, axes=(0, 1)) / N) < 1e-13
    X = _rand_c(rng, (N, bw + 1, N))
    back = s._irfft(X)
    assert rel_l2(back, np.fft.irfftn(X, s=(N, N), axes=(0, 1)) * N) < 1e-13
    assert rel_l2(s._irfft(got), f) < 1e-13


def test_bench_hook_copy_reference():
    """so3ft_bench_kernel id 13, the streaming copy that serves as the FFT passes' bandwidth
    reference: it copies T complex64 grids exactly and reports a positive duration."""
    import ctypes as C
    from pysofft_amd import _lib
    bw, T = 8, 3
    N = 2 * bw
    s = _soft(bw)
    a = torch.randn(T, N, N, N, dtype=torch.complex64, device="cuda")
    b = torch.zeros_like(a)
    out = C.c_float()
    _lib.check(_lib.lib().so3ft_bench_kernel(s._handle, 13, 2, C.c_void_p(a.data_ptr()), C.c_void_p(b.data_ptr()),
                                             T, C.c_void_p(torch.cuda.current_stream().cuda_stream), C.byref(out)))
    torch.cuda.synchronize()
    assert out.value > 0
    assert torch.equal(a, b)
