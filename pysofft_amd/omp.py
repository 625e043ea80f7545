"""Compatibility stub for ``pysofft.omp`` (pysofft/omp.py:7-32).

GPU transforms do not use OpenMP; the thread count is recorded and returned so
code written for the reference keeps working.
"""
import os

_n = max(1, (os.cpu_count() or 2) // 2)


def set_num_threads(n=None):
    global _n
    _n = max(1, int(n)) if n is not None else max(1, (os.cpu_count() or 2) // 2)


def get_max_threads():
    return _n
