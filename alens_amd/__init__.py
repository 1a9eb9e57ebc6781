"""MI355X-native collision-constraint engine for aLENS/SimToolbox rod systems.

The product is libalens.so (HIP kernels for gfx950 behind the C ABI in include/alens_capi.h);
this package holds its build script, the ctypes binding and the workload generators.
"""
from . import capi, systems  # noqa: F401
from .capi import AlensError, Context  # noqa: F401

__all__ = ["capi", "systems", "Context", "AlensError"]
