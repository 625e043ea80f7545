"""Compatibility stub for ``pysofft.fftw`` (pysofft/fftw.py:16-58).

The transforms here run on the GPU with the library's own LDS FFT kernels, so FFTW
planner flags and wisdom files have no effect.  The names are kept so that code
written for the reference keeps importing and calling them.
"""
from collections import namedtuple

flags = namedtuple('FFTWFlags', ['FFTW_MEASURE', 'FFTW_ESTIMATE'])(0, 64)
_wisdom_path = None


def set_wisdom_file_path(path):
    global _wisdom_path
    _wisdom_path = path


def load_wisdom_from_file(path=None):
    return False


def save_wisdom_to_file(path=None):
    return False
