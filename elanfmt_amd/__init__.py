"""MI355X-native staggered-grid elastodynamic FDTD time loop.

libfwe.so (gfx950 HIP kernels behind the C ABI of include/fwe.h) is the
product; this package holds its ctypes mirror (fwe.py), the workload
builders of the benchmark configurations (workloads.py), and the sources
(csrc/) with their in-tree build.
"""
from . import fwe  # noqa: F401

__all__ = ["fwe"]
