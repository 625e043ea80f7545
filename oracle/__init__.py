"""CPU oracle for the SO(3) Fourier transform path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may
import this package, and only as the checker (never as the thing measured or
shipped).  The product package ``pysofft_amd`` does not import it.

Contents
--------
``liboracle.so``   C restatement of the reference algorithm (so3_oracle.c), FP64.
``_ref/libsoftref.so``
                   the reference's own ``utils.f90``/``wigner.f90`` compiled in place
                   (oracle/Makefile); present only where it was built.
``brute``          numpy/scipy restatements from the closed-form definitions
                   (docs/defs/wigner.md:9-16), used to pin both of the above.

What pins each oracle function is tested in tests/test_oracle_pins.py.
"""
from .oracle import *  # noqa: F401,F403
from . import brute  # noqa: F401
