"""Zero-allocation forward transform so3ft_soft_many_c64_noalloc against so3ft_soft_many_c64.

Both are called through the C ABI on the same complex64 inputs; the zero-allocation entry
point runs the same kernels with its scratch in the library's static device array, in
passes of as many transforms as that array holds.  Two calls in a row, the second with a
smaller input, must each equal the allocating entry point's output (relative L2 <= 1e-7
per transform, the bar of the complex64 path; the kernels are the same, so the results
are expected to be identical).
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _call(fn, s, x, n):
    from pysofft_amd import _lib
    out = torch.full((n, s.n_coeffs), complex(np.nan, np.nan), dtype=torch.complex64, device="cuda")
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    _lib.check(fn(s._handle, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), n, stream))
    return out


def _check(s, n, seed):
    from pysofft_amd import _lib
    L = _lib.lib()
    N = 2 * s.bw
    g = torch.Generator(device="cuda").manual_seed(seed)
    x = torch.view_as_complex(torch.randn(n, N, N, N, 2, generator=g, device="cuda")).contiguous()
    got = _call(L.so3ft_soft_many_c64_noalloc, s, x, n)
    ref = _call(L.so3ft_soft_many_c64, s, x, n)
    torch.cuda.synchronize()
    assert torch.isfinite(torch.view_as_real(got)).all()
    for i in range(n):
        e = ((got[i].to(torch.complex128) - ref[i].to(torch.complex128)).norm() / ref[i].norm()).item()
        assert e <= 1e-7, (s.bw, n, i, e)
    return bool(torch.equal(got, ref))


@pytest.mark.parametrize("bw,n1,n2,rt", [(64, 20, 7, 0), (128, 3, 1, 0), (16, 5, 2, 1)])
def test_soft_many_c64_noalloc_matches_allocating_entry(bw, n1, n2, rt):
    """bw=64: 15 transforms per pass, so the first call takes two passes; bw=128: 2 per pass;
    bw=16 on a Risbo handle (lmn order, coefficient staging in the same array)."""
    from pysofft_amd import Soft
    s = Soft(bw, recurrence_type=rt)
    _check(s, n1, 10 * bw + n1)
    _check(s, n2, 10 * bw + n2)


def test_soft_many_c64_noalloc_rejects_oversize():
    from pysofft_amd import Soft, _lib
    s = Soft(256)
    x = torch.zeros(1, 512, 512, 512, dtype=torch.complex64, device="cuda")
    out = torch.zeros(1, s.n_coeffs, dtype=torch.complex64, device="cuda")
    rc = _lib.lib().so3ft_soft_many_c64_noalloc(s._handle, C.c_void_p(x.data_ptr()), C.c_void_p(out.data_ptr()), 1,
                                                C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1
    assert b"static" in _lib.lib().so3ft_last_error()
