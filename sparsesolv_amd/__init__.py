"""MI355X (gfx950) shifted-IC Krylov solvers for real and complex symmetric
sparse systems: shifted ICCG / ICCOCG, shifted-IC MRTR and Eisenstat SGS-MRTR,
behind the reference's solver interface.

The compute path is the HIP library lib/libsparsesolv_amd.so (C-ABI in
include/sparsesolv_amd.h); `capi` binds it with ctypes.  There is no CPU
fallback: without the built library every solve raises.
"""
from . import capi  # noqa: F401

__all__ = ["capi"]
