"""Host-side logic of pysofft_amd (no GPU needed): index utilities, the C ABI
library exports, and the failure behaviour without a GPU."""
import ctypes

import numpy as np
import pytest

from pysofft_amd import utils, wigner, _lib


@pytest.mark.parametrize("bw", [1, 2, 3, 5, 8, 13])
def test_coeff_degrees_vs_reference_vectors(golden, bw):
    assert np.array_equal(utils.get_coeff_degrees(bw), golden[f"degrees_{bw}"].astype(np.int64))
    deg = utils.get_coeff_degrees(bw)
    for i in range(0, len(deg), max(1, len(deg) // 50)):
        l, m, n = deg[i]
        assert utils.coeff_location_mnl(m, n, l, bw) - 1 == i


@pytest.mark.parametrize("bw", [1, 4, 6])
def test_host_enforce_real_sym_vs_reference_vectors(golden, bw):
    c = golden[f"sym_in_{bw}"].copy()
    utils.enforce_real_sym_mnl(c, bw)
    np.testing.assert_array_equal(c, golden[f"sym_out_{bw}"])


@pytest.mark.parametrize("bw", [1, 2, 7, 16, 64, 256])
def test_host_weights(golden, bw):
    np.testing.assert_allclose(utils.legendre_quadrature_weights(bw), golden[f"weights_{bw}"], rtol=0,
                               atol=4e-16 * bw)


def test_sampling_and_small_index_functions():
    np.testing.assert_allclose(wigner.create_beta_samples(10), (2 * np.arange(10) + 1) * np.pi / 20)
    np.testing.assert_allclose(wigner.create_alpha_gamma_samples(8), np.arange(8) * np.pi / 4)
    assert utils.total_num_coeffs(32) == 43680
    bw = 9
    seen = set()
    for l in range(bw):
        for m in range(-l, l + 1):
            seen.add(utils.mlc(m, l, bw))
            assert utils.lmc(l, m) == l * (l + 1) + m + 1
    assert len(seen) == bw * bw and min(seen) == 1 and max(seen) == bw * bw
    seen = {utils.mlr(m, l, bw) for m in range(bw) for l in range(m, bw)}
    assert seen == set(range(1, bw * (bw + 1) // 2 + 1))


def test_library_loads_and_exports_header_symbols():
    L = _lib.lib()
    names = _lib.header_symbols()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert L.so3ft_get_kostelec_recurrence_val() == 0
    assert L.so3ft_get_risbo_recurrence_val() == 1


def test_invalid_arguments_return_codes_not_aborts():
    L = _lib.lib()
    h = ctypes.c_int64()
    assert L.so3ft_init(0, 0, 0, 0, 0, 0, ctypes.byref(h)) == -1        # bw < 1
    assert L.so3ft_init(8, 8, 0, 0, 0, 0, ctypes.byref(h)) == -1        # lmax >= bw
    assert L.so3ft_init(8, 7, 0, 0, 5, 0, ctypes.byref(h)) == -1        # unknown recurrence
    assert b"recurrence" in L.so3ft_last_error()
    assert L.so3ft_destroy(123456) == -2
    assert L.so3ft_isoft(123456, None, None, None) == -2


def test_soft_fails_loudly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from pysofft_amd import Soft
    with pytest.raises(RuntimeError):
        Soft(8)


def test_group_layout_host():
    """Packed layout of a ragged batch: bandwidths in decreasing order, each bandwidth's
    transforms contiguous in the caller's order; unsupported bandwidths rejected."""
    from pysofft_amd.grouped import group_layout
    bws = [32, 128, 48, 32, 96, 64, 128, 48]
    co, go, ct, gt = group_layout(bws)
    nc = lambda b: (4 * b ** 3 - b) // 3
    order = sorted(range(len(bws)), key=lambda i: (-bws[i], i))
    c = g = 0
    for i in order:
        assert co[i] == c and go[i] == g, i
        c += nc(bws[i])
        g += (2 * bws[i]) ** 3
    assert (ct, gt) == (c, g)
    L = _lib.lib()
    b = np.array([32, 300], np.int64)
    assert L.so3ft_group_layout(b.ctypes.data, 2, None, None, None, None) == -5
    assert b"2 <= bw <= 256" in L.so3ft_last_error()
    b = np.array([1], np.int64)
    assert L.so3ft_group_layout(b.ctypes.data, 1, None, None, None, None) == -5
    assert L.so3ft_group_layout(b.ctypes.data, 0, None, None, None, None) == -1
    assert L.so3ft_group_destroy(99999) == -2


@pytest.mark.parametrize("bw", [1, 2, 3, 5, 8, 13])
def test_lmn_degrees_vs_reference_vectors(golden, bw):
    """lmn order (Risbo layout) against the reference's get_coeff_degrees_risbo, with
    coeff_location_lmn its inverse and lmn_to_mnl the reordering onto the mnl layout."""
    d = utils.get_coeff_degrees_risbo(bw)
    assert np.array_equal(d, golden[f"degrees_risbo_{bw}"].astype(np.int64))
    for i in range(0, len(d), max(1, len(d) // 60)):
        l, m, n = d[i]
        assert utils.coeff_location_lmn(l, m, n) - 1 == i
    p = utils.lmn_to_mnl(bw)
    assert len(np.unique(p)) == len(p) == utils.total_num_coeffs(bw)
    assert np.array_equal(utils.get_coeff_degrees(bw)[p], d)


@pytest.mark.parametrize("bw", [1, 4, 6])
def test_host_enforce_real_sym_lmn_vs_reference_vectors(golden, bw):
    c = golden[f"sym_in_{bw}"].copy()
    utils.enforce_real_sym_lmn(c, bw)
    np.testing.assert_array_equal(c, golden[f"sym_lmn_out_{bw}"])
