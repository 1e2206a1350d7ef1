"""mara3p_amd: Mara3's Godunov block update, MI355X-native.

The hot path (guard fill, PLM, HLLE/HLLC, source terms, RK update, CFL
reduction) runs in hand-written HIP kernels for gfx950 behind the C ABI in
include/mara3p_amd.h.  This package binds that ABI and keeps meshes resident
in device memory.  Importing it loads the HIP library and fails if it has not
been built -- there is no CPU fallback.
"""
from ._lib import (M3P_FP64, M3P_MIXED, M3PError, MinidiskData, NegativeDensityError, check,
                   check_kernel_errors, lib, LIBRARY_PATH)
from .mesh import ChainMesh1D, UniformMesh2D, to_device_layout, to_reference_layout

lib()

__all__ = [
    "M3P_FP64", "M3P_MIXED", "M3PError", "MinidiskData", "NegativeDensityError", "check",
    "check_kernel_errors", "lib", "LIBRARY_PATH", "ChainMesh1D", "UniformMesh2D",
    "to_device_layout", "to_reference_layout",
]
