"""ctypes binding of the C ABI in include/mara3p_amd.h.

The product path has exactly one implementation: the HIP library
``libmara3p_amd.so`` built next to this file.  If it is missing, importing
``mara3p_amd`` raises -- there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# The in-tree build; MARA3P_AMD_LIBRARY may name another build of the same library.
LIBRARY_PATH = os.environ.get("MARA3P_AMD_LIBRARY", os.path.join(_HERE, "libmara3p_amd.so"))

M3P_OK = 0
M3P_ERR_NEGATIVE_DENSITY = 1
M3P_ERR_NEGATIVE_PRESSURE = 2
M3P_ERR_NO_CONVERGENCE = 3
M3P_ERR_BAD_SHAPE = 4
M3P_ERR_HIP = 5
M3P_ERR_BAD_ARGUMENT = 6

M3P_FP64 = 0
M3P_MIXED = 1
M3P_KIND_EULER2D = 0
M3P_KIND_MINIDISK = 1
M3P_KIND_AMR_EULER2D = 2

PRECISIONS = {"fp64": M3P_FP64, "mixed": M3P_MIXED}


class M3PError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


class NegativeDensityError(M3PError, ValueError):
    """Raised where the reference throws std::invalid_argument
    ("euler::recover_primitive (negative density)", src/physics_euler.hpp:158-161)."""


class NegativePressureError(M3PError, ValueError):
    """Raised where the reference throws std::invalid_argument for a negative
    pressure (mhd::recover_primitive, src/physics_mhd.hpp:296-300)."""


class NoConvergenceError(M3PError, ValueError):
    """Raised where the reference throws std::invalid_argument for a root finder
    that does not converge (srhd::recover_primitive, src/physics_srhd.hpp:273-277)."""


class MinidiskData(ctypes.Structure):
    """problems/minidisk.hpp:80-96 solver_data_t (defaults :99-122)."""
    _fields_ = [
        ("block_size", ctypes.c_int),
        ("depth", ctypes.c_int),
        ("omega_frame", ctypes.c_double),
        ("mach_number", ctypes.c_double),
        ("kinematic_viscosity", ctypes.c_double),
        ("domain_radius", ctypes.c_double),
        ("softening_length", ctypes.c_double),
        ("eccentricity", ctypes.c_double),
        ("mass_ratio", ctypes.c_double),
        ("sink_radius", ctypes.c_double),
        ("sink_rate", ctypes.c_double),
        ("buffer_rate", ctypes.c_double),
        ("buffer_scale", ctypes.c_double),
        ("cfl", ctypes.c_double),
    ]

    @classmethod
    def defaults(cls, **kw):
        d = dict(block_size=64, depth=1, omega_frame=1.0, mach_number=10.0, kinematic_viscosity=0.0,
                 domain_radius=1.5, softening_length=0.01, eccentricity=0.0, mass_ratio=1.0,
                 sink_radius=0.01, sink_rate=1e3, buffer_rate=1e2, buffer_scale=0.2, cfl=0.3)
        d.update(kw)
        return cls(**d)

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_L = ctypes.c_long

# name -> (restype, argtypes); must match include/mara3p_amd.h exactly.
SIGNATURES = {
    "m3p_version": (_I, []),
    "m3p_last_error": (ctypes.c_char_p, []),
    "m3p_check_errors": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "m3p_euler2d_updated_conserved": (_I, [_P, _P, _P, _I, _D, _D, _D, _D, _I, _I, _P]),
    "m3p_euler1d_updated_conserved": (_I, [_P, _P, _P, _I, _D, _D, _D, _D, _I, _I, _I, _P]),
    "m3p_minidisk_updated_conserved": (_I, [_P, _P, _P, _I, _I, _I, _D, _D, ctypes.POINTER(MinidiskData), _I, _P]),
    "m3p_locally_isothermal_updated_conserved": (_I, [_P, _P, _P, _I, _D, _D, _P, _P, _D, _I, _P]),
    "m3p_locally_isothermal_stage": (_I, [_P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _P, _P, _D, _I, _P]),
    "m3p_euler2d_stage": (_I, [_P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _I, _I, _P]),
    "m3p_minidisk_stage": (_I, [_P, _P, _P, _P, _P, _I, _D, _D, _D, ctypes.POINTER(MinidiskData), _I, _P]),
    "m3p_minidisk_clock": (_I, [ctypes.POINTER(MinidiskData), _P, _P, _I, _P]),
    "m3p_minidisk_stage_dev": (_I, [_P, _P, _P, _P, _P, _I, _P, _P, _D, ctypes.POINTER(MinidiskData), _I, _P]),
    "m3p_minidisk_rk2_step": (_I, [_P, _P, _P, _P, _I, _D, _D, _D, ctypes.POINTER(MinidiskData), _I, _P]),
    "m3p_minidisk_rk2_step_dev": (_I, [_P, _P, _P, _P, _I, _P, _P, _D, ctypes.POINTER(MinidiskData), _I, _P]),
    "m3p_euler2d_rk2_step": (_I, [_P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _I, _I, _P]),
    "m3p_euler2d_rk2_step_cfl": (_I, [_P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _I, _D, _P, _I, _P]),
    "m3p_euler1d_rk2_step_blocks": (_I, [_P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _I, _P]),
    "m3p_euler1d_rk2_step_blocks_cfl": (_I, [_P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _P, _I, _P]),
    "m3p_euler1d_stage": (_I, [_P, _P, _P, _P, _I, _I, _D, _D, _D, _D, _D, _I, _P]),
    "m3p_euler_cfl_dt": (_I, [_P, _I, _L, _D, _D, _D, _P, _P]),
    "m3p_set_stage2d_tile": (_I, [_I, _I, _I, _I]),
    "m3p_get_stage2d_tile": (_I, [_I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "m3p_autotune_euler2d": (_I, [_P, _P, _P, _I, _I, _D, _D, _I, _I, ctypes.POINTER(_I), ctypes.POINTER(_I),
                                  ctypes.POINTER(ctypes.c_float), _P]),
    "m3p_autotune_minidisk": (_I, [_P, _P, _P, _P, _I, ctypes.POINTER(MinidiskData), _I, _I, ctypes.POINTER(_I),
                                   ctypes.POINTER(_I), ctypes.POINTER(ctypes.c_float), _P]),
    "m3p_conserved_totals": (_I, [_P, _I, _I, _L, _P, _P]),
    "m3p_euler_recover_primitive_array": (_I, [_P, _P, _L, _D, _P]),
    "m3p_iso2d_recover_primitive_array": (_I, [_P, _P, _L, _P]),
    "m3p_plm_gradient_2d": (_I, [_P, _P, _I, _I, _I, _I, _D, _P]),
    "m3p_weighted_sum": (_I, [_P, _P, _P, _L, _D, _P]),
    "m3p_mhd_workspace_bytes": (_L, [_I, _I]),
    "m3p_mhd_advance_ct": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _I, _D, _D, _D, _D, _P, _P, _P, _P, _P,
                                _P, _L, _I, _P]),
    "m3p_mhd_advance_ct_noalloc": (_I, [_P, _P, _P, _P, _P, _P, _P, _P, _D, _I, _D, _D, _D, _D, _P, _P, _P, _P, _P,
                                        _P, _I, _P]),
    "m3p_mhd_max_div_b": (_I, [_P, _P, _P, _I, _P, _P]),
    "m3p_mhd_v2_primitive_array": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P]),
    "m3p_mhd_v2_godunov_fluxes": (_I, [_P, _P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "m3p_mhd_v2_electromotive_forces": (_I, [_P, _P, _P, _I, _I, _I, _P, _P, _P, _P]),
    "m3p_mhd_v2_updated_conserved_density": (_I, [_P, _P, _P, _P, _I, _I, _I, _D, _D, _P, _P]),
    "m3p_mhd_v2_curl": (_I, [_P, _P, _P, _I, _I, _I, _D, _P, _P, _P, _P]),
    "m3p_mhd_v2_updated_magnetic_flux_density": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _I, _D, _D, _P, _P, _P, _P]),
    "m3p_mhd_v2_extend_cell_primitive_variables": (_I, [_P, _I, _I, _P, _P]),
    "m3p_mhd_v2_extend_face_magnetic_flux_density": (_I, [_P, _I, _I, _P, _P, _P, _P]),
    "m3p_srhd_recover_primitive_array": (_I, [_P, _P, _L, _D, _D, _P]),
    "m3p_euler1d_stage_blocks": (_I, [_P, _P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _I, _P]),
    "m3p_amr_euler2d_stage": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _I, _I, _P]),
    "m3p_amr_euler2d_stage_conservative": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _I, _I, _P]),
    "m3p_euler_cfl_dt_blocks": (_I, [_P, _I, _L, _P, _D, _D, _P, _P]),
    "m3p_euler2d_stage_cfl": (_I, [_P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _I, _D, _P, _I, _P]),
    "m3p_euler1d_stage_blocks_cfl": (_I, [_P, _P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _D, _P, _I, _P]),
    "m3p_amr_euler2d_stage_cfl": (_I, [_P, _P, _P, _P, _P, _P, _I, _I, _D, _P, _D, _D, _D, _I, _D, _P, _I, _P]),
    "m3p_graph_begin": (_I, [_P]),
    "m3p_graph_end": (_I, [_P, ctypes.POINTER(_P)]),
    "m3p_graph_launch": (_I, [_P, _P]),
    "m3p_graph_destroy": (_I, [_P]),
    "m3p_srhd_conserved_density_array": (_I, [_P, _P, _L, _D, _P]),
    "m3p_srhd_riemann_array": (_I, [_P, _P, _P, _L, _I, _D, _I, _D, _P]),
}

_lib = None


def lib():
    """Load (once) and return the HIP library; raise if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBRARY_PATH):
            raise ImportError(
                "mara3p_amd: HIP library %s is not built (run python mara3p_amd/build.py, "
                "which compiles it for gfx950 in-tree)" % LIBRARY_PATH)
        # The library shares device pointers and streams with PyTorch, so it
        # must bind to the same HIP runtime.  PyTorch ships its own
        # libamdhip64.so.7; loading torch first makes that the process's copy
        # (same soname), instead of the system one this library was linked
        # against being loaded first and then picked up by torch as well.
        import torch  # noqa: F401
        handle = ctypes.CDLL(LIBRARY_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def last_error():
    return lib().m3p_last_error().decode()


def check(status):
    """Turn an m3p status code into an exception (reference error behaviour)."""
    if status == M3P_OK:
        return
    msg = last_error()
    if status == M3P_ERR_NEGATIVE_DENSITY:
        raise NegativeDensityError(status, msg)
    if status == M3P_ERR_NEGATIVE_PRESSURE:
        raise NegativePressureError(status, msg)
    if status == M3P_ERR_NO_CONVERGENCE:
        raise NoConvergenceError(status, msg)
    raise M3PError(status, msg)


def check_kernel_errors(stream=0):
    """Synchronise and raise if a kernel latched an error (e.g. negative density)."""
    code, block, zone = _I(0), _I(0), _I(0)
    check(lib().m3p_check_errors(_P(stream), ctypes.byref(code), ctypes.byref(block), ctypes.byref(zone)))
    if code.value == M3P_ERR_NEGATIVE_DENSITY:
        raise NegativeDensityError(code.value, "euler::recover_primitive (negative density) in block %d" % block.value)
    if code.value == M3P_ERR_BAD_SHAPE:
        raise M3PError(code.value, "block %d: its prev/next links are not mutual (the fused 1D RK2 step needs them "
                                   "to be)" % block.value)
    if code.value:
        raise M3PError(code.value, "kernel error %d in block %d" % (code.value, block.value))
