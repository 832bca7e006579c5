"""MI355X (gfx950) implicit-solve path of DRD-Micro.

The product is the C-ABI library libdrd_gpu.so (include/drd_gpu.h, HIP
sources in csrc/) and the Fortran drop-in shim (fortran/drd_gpu_shim.f90).
This package is the Python host interface over the same C-ABI.
"""
from . import abi  # noqa: F401

__all__ = ["abi", "solver", "geometry"]
