"""spectral2d_amd -- MI355X-native TVD-RK3 time step of mpclay/Spectral2D.

The compute path is libs2d.so (hand-written HIP for gfx950, C ABI in
include/s2d.h). This package holds the build recipe, its ctypes binding and
a Python mirror of the reference's Spectral_m / TimeIntegration_m interfaces.
The Fortran drop-in (ISO_C_BINDING) lives in spectral2d_amd/fortran.
"""
from ._lib import (DEALIAS_3_2_RULE, FP32, FP64, NO_DEALIASING, RK3_TVD_SHU, S2DError,
                   LIB_PATH)

__all__ = ["DEALIAS_3_2_RULE", "NO_DEALIASING", "RK3_TVD_SHU", "FP32", "FP64", "S2DError",
           "LIB_PATH"]
