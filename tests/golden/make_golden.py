"""Writes tests/golden/ref_vectors.npz from the reference's own Fortran modules
(utils.f90 / wigner.f90 compiled in place by oracle/Makefile into oracle/_ref).

Run from the repository root in the container that holds the reference checkout:
    make -C oracle all && python tests/golden/make_golden.py

Every array is data produced by the reference's code (index layouts, weights,
Wigner-d seeds and matrices, the real-symmetry fill) or a direct sum over the
reference's own d-matrices (isoft values); the CPU tests check the oracle and the
host utilities against these vectors without needing the reference build.
"""
import os
import sys

import numpy as np

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", ".."))
sys.path.insert(0, ROOT)

from oracle import RefLib as R  # noqa: E402


def main():
    assert R.available(), "oracle/_ref/libsoftref.so is not built"
    out = {}
    rng = np.random.default_rng(20240611)
    for bw in [1, 2, 3, 5, 8, 13]:
        out[f"degrees_{bw}"] = R.coeff_degrees(bw).astype(np.int16)
        out[f"degrees_risbo_{bw}"] = R.coeff_degrees_risbo(bw).astype(np.int16)
    for bw in [1, 2, 3, 4, 7, 16, 32, 64, 128, 256]:
        out[f"weights_{bw}"] = R.legendre_weights(bw)
    out["seeds_9"] = R.seeds(9)
    for (p, q, bw) in [(0, 0, 16), (3, 7, 16), (5, 5, 16), (0, 15, 16), (32, 56, 129), (0, 0, 128), (17, 90, 128),
                       (127, 127, 128)]:
        out[f"genwig_{p}_{q}_{bw}"] = R.genwig(p, q, bw)
    betas = np.array([np.pi * 5 / 11, np.pi - np.pi * 5 / 11])
    for l in [2, 7, 30]:
        out[f"dl_kostelec_{l}"] = R.wigner_dl_kostelec(l, betas, False)
    out["dl_betas"] = betas
    # Risbo steps (unnormalised full small-d, wigner.f90:529-585) at one beta
    sq = np.sqrt(np.arange(2 * 12 + 1))
    d = None
    cb, sb = np.cos(0.83), np.sin(-0.83)
    for l in range(12):
        d = R.risbo_step(d, l, cb, sb, sq)
    out["risbo_l11"] = d
    out["risbo_cs"] = np.array([cb, sb])
    # real-symmetry fill
    for bw in [1, 4, 6]:
        nc = (4 * bw ** 3 - bw) // 3
        c = rng.standard_normal(nc) + 1j * rng.standard_normal(nc)
        out[f"sym_in_{bw}"] = c.copy()
        out[f"sym_out_{bw}"] = R.enforce_real_sym_mnl(c, bw)
        out[f"sym_lmn_out_{bw}"] = R.enforce_real_sym_lmn(out[f"sym_in_{bw}"], bw)
    # isoft on the grid from the reference's normalised d matrices (wigner_dl_kostelec):
    # f(g_j, a_i, b_k) = sum c_lmn e^{-i m a} d~^l_{mn}(b) e^{-i n g} / (2 pi)
    for bw in [2, 3, 4, 8]:
        N = 2 * bw
        deg = R.coeff_degrees(bw)
        c = rng.standard_normal(len(deg)) + 1j * rng.standard_normal(len(deg))
        a = R.alpha_gamma_samples(N)
        b = R.beta_samples(N)
        dls = {l: R.wigner_dl_kostelec(l, b, True) for l in range(bw)}
        ea = np.exp(-1j * np.outer(np.arange(-bw + 1, bw), a))
        f = np.zeros((N, N, N), np.complex128)
        for i, (l, m, n) in enumerate(deg):
            dv = dls[l][:, n + l, m + l]
            f += c[i] * (ea[n + bw - 1][:, None, None] * ea[m + bw - 1][None, :, None] * dv[None, None, :])
        out[f"isoft_c_{bw}"] = c
        out[f"isoft_f_{bw}"] = f / (2 * np.pi)
    # the symmetry-reduced Risbo recurrence (wigner.f90:626-748) through genwig_all_risbo
    # (wigner.f90:784-822, mnl rows x 2bw betas) and the full Wigner-D of full_big_d_risbo
    # (wigner.f90:844-934); the Kostelec table (wigner.f90:423-465) for comparison
    out["genwig_all_risbo_9"] = R.genwig_all_risbo(9, True)
    out["genwig_all_risbo_raw_5"] = R.genwig_all_risbo(5, False)
    out["genwig_all_kostelec_9"] = R.genwig_all_kostelec(9)
    out["full_big_d_angles"] = np.array([0.3, 1.1, 2.5])
    out["full_big_d_6"] = R.full_big_d_norm(6, 0.3, 1.1, 2.5, True)
    out["full_big_d_raw_4"] = R.full_big_d_norm(4, 0.3, 1.1, 2.5, False)
    path = os.path.join(os.path.dirname(__file__), "ref_vectors.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
