"""GPU tests of the learned alignment proposal and the local correlation refinement.

The network's kernels (BF16 operands, FP32 accumulation) are checked against the FP64
oracle in oracle/mlp_oracle.py (itself pinned by finite differences); the refinement is
checked against the full correlation grid, which the transform tests pin to the oracle.
"""
import math

import numpy as np
import pytest

import oracle as O
from conftest import rel_l2

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _bf16_round(a):
    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).to(torch.float32).numpy()


def test_gemm_bf16_vs_numpy():
    import ctypes as C
    from pysofft_amd import _lib
    rng = np.random.default_rng(0)
    for (M, N, K) in [(70, 45, 37), (128, 64, 16), (5, 130, 200)]:
        A = torch.from_numpy(rng.standard_normal((M, K)).astype(np.float32)).to(torch.bfloat16).cuda()
        B = torch.from_numpy(rng.standard_normal((K, N)).astype(np.float32)).to(torch.bfloat16).cuda()
        Cm = torch.empty((M, N), dtype=torch.float32, device="cuda")
        _lib.check(_lib.lib().so3ft_gemm_bf16(C.c_void_p(A.data_ptr()), C.c_void_p(B.data_ptr()),
                                              C.c_void_p(Cm.data_ptr()), M, N, K, None))
        ref = A.float().cpu().numpy().astype(np.float64) @ B.float().cpu().numpy().astype(np.float64)
        assert rel_l2(Cm.cpu().numpy(), ref) < 1e-6, (M, N, K)


def _random_rotations(rng, n):
    Q, R = np.linalg.qr(rng.standard_normal((n, 3, 3)))
    Q = Q * np.sign(np.diagonal(R, axis1=1, axis2=2))[:, None, :]
    Q[np.linalg.det(Q) < 0, :, 2] *= -1
    return Q


def test_mlp_forward_and_gradients_vs_oracle():
    """The network's forward pass, loss and per-layer gradients against the FP64 oracle that
    rounds to bfloat16 where the device does (oracle/mlp_oracle.py, itself pinned by finite
    differences): relative L2 <= 1e-2 (what is left is FP32 accumulation and the rounding
    of values it moves across a bfloat16 boundary); the unrounded oracle is within 3e-2.
    Then one Adam step (lr > 0) against a numpy Adam on the device's own gradients."""
    from oracle import mlp_oracle as M
    from pysofft_amd.align import AlignmentMLP, init_params
    rng = np.random.default_rng(1)
    dims = (452, 96, 80, 64, 6)
    n = 50
    p = init_params(dims, seed=3)
    net = AlignmentMLP(dims, params=p)
    X = rng.standard_normal((n, dims[0])).astype(np.float32)
    T = _random_rotations(rng, n)
    Xt = torch.from_numpy(X).cuda()
    R, out = net.forward(Xt)
    o_ref, _, _ = M.forward_bf16(p, dims, X)
    assert rel_l2(out.cpu().numpy(), o_ref) < 1e-2
    assert rel_l2(R.cpu().numpy(), M.gram_schmidt(o_ref)) < 1e-2
    assert rel_l2(out.cpu().numpy(), M.forward(p, dims, X)[0]) < 3e-2
    Tt = torch.from_numpy(T.astype(np.float32)).cuda()
    loss = net.train_step(Xt, Tt, 0.0)
    l_ref = np.mean(np.sum((M.gram_schmidt(o_ref) - T) ** 2, axis=(1, 2)))
    assert abs(loss - l_ref) < 1e-2 * abs(l_ref)
    g = net.get_params(grads=True)
    g_ref = M.gradients_bf16(p, dims, X, T.reshape(n, 9))
    o = 0
    for i in range(len(dims) - 1):
        nw, nbias = dims[i] * dims[i + 1], dims[i + 1]
        assert rel_l2(g[o:o + nw], g_ref[o:o + nw]) < 1e-2, ("W", i, rel_l2(g[o:o + nw], g_ref[o:o + nw]))
        o += nw
        assert rel_l2(g[o:o + nbias], g_ref[o:o + nbias]) < 1e-2, ("b", i, rel_l2(g[o:o + nbias], g_ref[o:o + nbias]))
        o += nbias
    # lr = 0 leaves the weights unchanged
    np.testing.assert_array_equal(net.get_params(), p)
    # one Adam step (the first: zero moments, bias corrections 1 - 0.9, 1 - 0.999)
    lr = 1e-3
    net.train_step(Xt, Tt, lr)
    g1 = net.get_params(grads=True).astype(np.float64)
    w1 = net.get_params().astype(np.float64)
    w_ref, _, _ = M.adam_step(p.astype(np.float64), g1, 0.0, 0.0, 1, lr)
    np.testing.assert_allclose(w1, w_ref, rtol=0, atol=1e-6 * lr + 2e-7 * np.abs(w_ref).max())
    assert np.abs(w1 - p).max() > 0.5 * lr   # the step moved the weights


def test_features_vs_numpy():
    from pysofft_amd.align import features
    rng = np.random.default_rng(2)
    bw, n = 16, 3
    f = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    g = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    x = features(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda()).cpu().numpy()
    l = np.floor(np.sqrt(np.arange(bw * bw))).astype(int)
    for s in range(n):
        sc = np.sqrt(np.sum(np.abs(f[s, :81]) ** 2) / 81)
        ref = np.concatenate([f[s, :81].real / sc, f[s, :81].imag / sc, g[s, :81].real / sc, g[s, :81].imag / sc,
                              [np.log(np.sum(np.abs(f[s, l == k]) ** 2) / sc ** 2 + 1e-30) for k in range(64)],
                              [np.log(np.sum(np.abs(g[s, l == k]) ** 2) / sc ** 2 + 1e-30) for k in range(64)]])
        np.testing.assert_allclose(x[s], ref.astype(np.float32), rtol=2e-6, atol=2e-5)


def test_refine_window_matches_full_correlation_grid():
    """One refinement level at the correlation's own grid returns the max |C| over the
    5 x 5 x 5 window of the full grid around the start point, at its location."""
    from pysofft_amd import Soft
    from pysofft_amd.align import Aligner, AlignmentMLP
    bw, n = 16, 6
    rng = np.random.default_rng(3)
    f = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    g = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    s = Soft(bw)
    cc = np.asarray(s.cross_correlation_ylm_cmplx_many(f, g))  # (n, gamma, alpha, beta)
    N = 2 * bw
    a = np.arange(N) * np.pi / bw
    b = (np.arange(N) + 0.5) * np.pi / N
    start = rng.integers(0, N, (n, 3))
    start[:, 1] = np.clip(start[:, 1], 2, N - 3)
    e0 = np.stack([a[start[:, 0]], b[start[:, 1]], a[start[:, 2]]], 1)
    al = Aligner(mlp=AlignmentMLP((452, 16, 6)))
    e, score = al.refine(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(), torch.from_numpy(e0).cuda(),
                         levels=[(bw, bw)])
    e, score = e.cpu().numpy(), score.cpu().numpy()
    for i in range(n):
        ia = (start[i, 0] + np.arange(-2, 3)) % N
        ib = start[i, 1] + np.arange(-2, 3)
        ig = (start[i, 2] + np.arange(-2, 3)) % N
        win = np.abs(cc[i][np.ix_(ig, ia, ib)])
        assert abs(score[i] - win.max()) < 1e-10 * win.max(), i
        jg, ja, jb = np.unravel_index(np.argmax(win), win.shape)
        np.testing.assert_allclose(e[i], [a[ia[ja]], b[ib[jb]], a[ig[jg]]], atol=1e-12)


def test_refine_window_on_a_finer_grid_matches_padded_correlation():
    """A refinement level whose grid bandwidth exceeds the sets' (grid 32, degrees < 16, the
    shape of the last ladder level) returns the max |C| over the 5 x 5 x 5 window of the
    bandwidth-32 correlation grid of the zero-padded sets, at its location."""
    from pysofft_amd import Soft
    from pysofft_amd.align import Aligner, AlignmentMLP
    bw, gb, n = 16, 32, 5
    rng = np.random.default_rng(31)
    f = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    g = rng.standard_normal((n, bw * bw)) + 1j * rng.standard_normal((n, bw * bw))
    fp = np.zeros((n, gb * gb), complex)
    gp = np.zeros((n, gb * gb), complex)
    fp[:, :bw * bw] = f   # LMc order: degrees < bw come first
    gp[:, :bw * bw] = g
    cc = np.asarray(Soft(gb).cross_correlation_ylm_cmplx_many(fp, gp))
    N = 2 * gb
    a = np.arange(N) * np.pi / gb
    b = (np.arange(N) + 0.5) * np.pi / N
    start = rng.integers(0, N, (n, 3))
    start[:, 1] = np.clip(start[:, 1], 2, N - 3)
    e0 = np.stack([a[start[:, 0]], b[start[:, 1]], a[start[:, 2]]], 1)
    al = Aligner(mlp=AlignmentMLP((452, 16, 6)))
    e, score = al.refine(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda(), torch.from_numpy(e0).cuda(),
                         levels=[(gb, bw)])
    e, score = e.cpu().numpy(), score.cpu().numpy()
    for i in range(n):
        ia = (start[i, 0] + np.arange(-2, 3)) % N
        ib = start[i, 1] + np.arange(-2, 3)
        ig = (start[i, 2] + np.arange(-2, 3)) % N
        win = np.abs(cc[i][np.ix_(ig, ia, ib)])
        assert abs(score[i] - win.max()) < 1e-10 * win.max(), i
        jg, ja, jb = np.unravel_index(np.argmax(win), win.shape)
        np.testing.assert_allclose(e[i], [a[ia[ja]], b[ib[jb]], a[ig[jg]]], atol=1e-12)


def test_training_reduces_loss_and_beats_untrained_proposals():
    """A short training run on generated pairs lowers the loss and the proposals' angular
    error against the untrained network (what the learned part gains over none)."""
    from pysofft_amd.align import Aligner, AlignmentMLP, training_batch
    from pysofft_amd import workloads as W
    al = Aligner(mlp=AlignmentMLP(seed=5))
    al.fit_normalisation(n=16384)
    gen = torch.Generator(device="cuda").manual_seed(77)
    f, g, T, e = training_batch(1024, gen)
    R0 = W.euler_to_matrix(al.propose(f, g).cpu().numpy())
    err0 = np.median(W.angular_distance(R0, T.cpu().numpy().astype(np.float64)))
    hist = al.train(pairs=4096 * 800, batch=4096, lr=1e-3)
    assert np.mean(hist[-5:]) < 0.8 * np.mean(hist[:5])
    R1 = W.euler_to_matrix(al.propose(f, g).cpu().numpy())
    err1 = np.median(W.angular_distance(R1, T.cpu().numpy().astype(np.float64)))
    assert err1 < 0.8 * err0, (np.degrees(err0), np.degrees(err1))


def test_align_end_to_end_within_one_grid_step():
    """Proposal + refinement + fallback on rotated noisy pairs at bw=64 (SNR 10): every
    rotation within one bw=64 grid step of the truth."""
    from pysofft_amd.align import Aligner, AlignmentMLP
    from pysofft_amd import workloads as W
    rng = np.random.default_rng(8)
    f, g, e = W.rotated_pairs(48, 64, rng, snr=10.0)
    al = Aligner(mlp=AlignmentMLP(seed=6))
    al.fit_normalisation(n=8192)
    al.train(pairs=4096 * 500, batch=4096)
    est, fell_back = al.align(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda())
    err = W.angular_distance(W.euler_to_matrix(est.cpu().numpy()), W.euler_to_matrix(e))
    assert np.all(err <= math.pi / 64), np.degrees(err.max())
    # most pairs are resolved by proposal + refinement, without the full correlation
    assert fell_back.float().mean().item() < 0.5


def test_trained_network_beats_network_free_proposal():
    """What the trained network gains over the method without it: trained with the default
    budget, its proposals on 2048 rotated noisy pairs at bw=64 (SNR 10) have a smaller median
    angular error than the network-free proposal (the peak of the correlation of the degrees
    below 4 on its 8^3 grid), and both pipelines put every pair within one bw=64 grid step."""
    from pysofft_amd.align import Aligner
    from pysofft_amd import workloads as W
    rng = np.random.default_rng(31)
    f, g, e = W.rotated_pairs(2048, 64, rng, snr=10.0)
    ft, gt = torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda()
    Rt = W.euler_to_matrix(e)
    al = Aligner()
    al.fit_normalisation()
    al.train()
    ab = Aligner(proposer="grid", threshold=al.threshold)

    def median_deg(A):
        return np.degrees(np.median(W.angular_distance(W.euler_to_matrix(A.propose(ft, gt).cpu().numpy()), Rt)))

    learned, grid = median_deg(al), median_deg(ab)
    assert learned < 0.8 * grid, (learned, grid)
    for A in (al, ab):
        est, _ = A.align(ft, gt)
        err = W.angular_distance(W.euler_to_matrix(est.cpu().numpy()), Rt)
        assert np.all(err <= math.pi / 64), np.degrees(err.max())


def test_align_fallback_search_within_one_grid_step():
    """The fallback alone (threshold above every normalised peak, so that every pair takes
    it): the full FP64 correlation at bw and its peak, then the last refinement window
    (grid bandwidth 2 bw).  Every rotation within one bw=64 grid step of the truth; the
    refined estimate beats the bw=64 grid's half step on most pairs."""
    from pysofft_amd.align import Aligner, AlignmentMLP
    from pysofft_amd import workloads as W
    rng = np.random.default_rng(21)
    f, g, e = W.rotated_pairs(24, 64, rng, snr=10.0)
    al = Aligner(mlp=AlignmentMLP(seed=3), threshold=2.0)   # no network training needed
    al.fit_normalisation(n=4096)
    est, fell_back = al.align(torch.from_numpy(f).cuda(), torch.from_numpy(g).cuda())
    assert bool(fell_back.all())
    err = W.angular_distance(W.euler_to_matrix(est.cpu().numpy()), W.euler_to_matrix(e))
    assert np.all(err <= math.pi / 64), np.degrees(err.max())
    assert np.median(err) < math.pi / 128, np.degrees(np.median(err))
