"""pysofft_amd -- SO(3) Fourier transforms on AMD Instinct MI355X (gfx950).

A drop-in for PySOFFT's ``Soft`` API whose transforms run as hand-written HIP
kernels in ``libso3ft.so`` (C ABI: include/so3ft.h).
"""
from . import fftw, omp, utils, wigner  # noqa: F401
from .soft import CoeffSO3, CoeffView, Soft  # noqa: F401

__all__ = ["Soft", "CoeffSO3", "CoeffView", "fftw", "omp", "utils", "wigner"]
