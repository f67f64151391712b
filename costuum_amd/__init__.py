"""MI355X-native (gfx950) T-matrix hot path of CosTuuM.

The numerics run in libcostuum_amd.so (HIP kernels behind the C ABI of
include/costuum_amd.h); this package holds the Python mirror of the
reference interface over that ABI.
"""
from . import _native  # noqa: F401

__all__ = ["_native"]
