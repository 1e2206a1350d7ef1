"""Python handle on the oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/liboracle.so (the C restatement) and, when built,
oracle/_ref/libmara_ref.so (the reference's own hot-path sources compiled
here).  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py may import this module; the product path never does.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmara_ref.so")

_D = ctypes.c_double
_I = ctypes.c_int
_L = ctypes.c_long
_P = ctypes.c_void_p


class OrcMinidisk(ctypes.Structure):
    _fields_ = [("block_size", _I), ("depth", _I)] + [(n, _D) for n in (
        "omega_frame", "mach_number", "kinematic_viscosity", "domain_radius", "softening_length",
        "eccentricity", "mass_ratio", "sink_radius", "sink_rate", "buffer_rate", "buffer_scale", "cfl")]

    @classmethod
    def from_data(cls, d):
        return cls(**{k: getattr(d, k) for k, _ in cls._fields_})

    def as_array(self):
        return np.array([getattr(self, k) for k, _ in self._fields_], dtype=np.float64)


_SIG = {
    "orc_plm": (_D, [_D, _D, _D, _D]),
    "orc_euler_recover_primitive": (_I, [_P, _D, _P]),
    "orc_euler_conserved_density": (None, [_P, _D, _P]),
    "orc_euler_flux": (None, [_P, _I, _D, _P]),
    "orc_euler_riemann_hlle": (None, [_P, _P, _I, _D, _P]),
    "orc_euler_riemann_hllc": (None, [_P, _P, _I, _D, _P]),
    "orc_iso2d_recover_primitive": (None, [_P, _P]),
    "orc_iso2d_conserved_density": (None, [_P, _P]),
    "orc_iso2d_riemann_hlle": (None, [_P, _P, _D, _I, _P]),
    "orc_two_body_state": (_I, [_D, _D, _D, _D, _D, _D, _P]),
    "orc_two_body_acceleration": (None, [_D, _D, _D, _D, _D, _D, _P]),
    "orc_euler1d_updated_conserved": (_I, [_P, _P, _I, _D, _D, _D, _D, _I, _I, _P]),
    "orc_euler2d_updated_conserved": (_I, [_P, _P, _I, _I, _D, _D, _D, _D, _I, _P]),
    "orc_locally_isothermal_updated_conserved": (_I, [_P, _P, _I, _I, _D, _D, _P, _P, _D, _P]),
    "orc_minidisk_updated_conserved": (_I, [_P, _P, ctypes.POINTER(OrcMinidisk), _I, _L, _L, _D, _D, _P]),
    "orc_minidisk_initial_conserved": (None, [ctypes.POINTER(OrcMinidisk), _I, _L, _L, _P]),
    "orc_minidisk_cell_coordinates": (None, [ctypes.POINTER(OrcMinidisk), _I, _L, _L, _P, _P]),
    "orc_extend_block_2d": (None, [_P, _I, _I, _I, _L, _L, _P]),
    "orc_extend_block_1d": (None, [_P, _I, _I, _I, _L, _P]),
    "orc_euler1d_mesh_rk2": (_I, [_P, _I, _I, _D, _D, _D, _D, _I, _I]),
    "orc_euler2d_mesh_rk2": (_I, [_P, _I, _I, _D, _D, _D, _D, _I, _I, _I]),
    "orc_euler2d_mesh_stage": (_I, [_P, _P, _I, _I, _D, _D, _D, _D, _I, _I]),
    "orc_minidisk_mesh_rk2": (_I, [_P, ctypes.POINTER(OrcMinidisk), _D, _D, _I, _I]),
    "orc_minidisk_mesh_stage": (_I, [_P, _P, ctypes.POINTER(OrcMinidisk), _D, _D, _I]),
    "orc_euler_exact_riemann": (None, [_P, _P, _D, _D, _P]),
    "orc_mhd_recover_primitive": (_I, [_P, _P, _D, _P]),
    "orc_mhd_conserved_density": (None, [_P, _D, _P]),
    "orc_mhd_flux": (None, [_P, _I, _D, _P]),
    "orc_mhd_fastest_speed": (_D, [_P, _D]),
    "orc_mhd_riemann_hlle": (None, [_P, _P, _D, _I, _D, _P, _P]),
    "orc_mhd_advance_ct": (_I, [_I, _P, _P, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P, _P]),
    "orc_mhd_rk2": (_I, [_I, _P, _P, _P, _P, _D, _D, _D, _D, _P]),
    "orc_mhd_v2_primitive_array": (_I, [_I, _I, _I, _P, _P, _P, _P, _P]),
    "orc_mhd_v2_godunov_fluxes": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
    "orc_mhd_v2_electromotive_forces": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "orc_mhd_v2_updated_conserved_density": (None, [_I, _I, _I, _P, _P, _P, _P, _D, _D, _P]),
    "orc_mhd_v2_curl": (None, [_I, _I, _I, _P, _P, _P, _D, _P, _P, _P]),
    "orc_mhd_v2_updated_magnetic_flux_density": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _D, _D, _P, _P, _P]),
    "orc_mhd_v2_extend_cells": (None, [_P, _I, _I, _I, _P]),
    "orc_mhd_v2_extend_faces": (None, [_P, _I, _I, _P, _P, _P]),
    "orc_srhd_recover_primitive": (_I, [_P, _D, _D, _P]),
    "orc_srhd_conserved_density": (None, [_P, _D, _P]),
    "orc_srhd_flux": (None, [_P, _I, _D, _P]),
    "orc_srhd_riemann": (_I, [_P, _P, _I, _D, _I, _D, _P]),
}

_REF_SIG = {
    "ref_plm": (_D, [_D, _D, _D, _D]),
    "ref_euler_recover_primitive": (_I, [_P, _D, _P]),
    "ref_euler_conserved_density": (None, [_P, _D, _P]),
    "ref_euler_flux": (None, [_P, _I, _D, _P]),
    "ref_euler_riemann_hlle": (None, [_P, _P, _I, _D, _P]),
    "ref_iso2d_recover_primitive": (None, [_P, _P]),
    "ref_iso2d_riemann_hlle": (None, [_P, _P, _D, _I, _P]),
    "ref_two_body_state": (_I, [_D, _D, _D, _D, _D, _D, _P]),
    "ref_two_body_acceleration": (None, [_D, _D, _D, _D, _D, _D, _P]),
    "ref_two_body_recover_elements": (None, [_P, _D, _P]),
    "ref_minidisk_updated_conserved": (None, [_P, _P, _P, _I, _L, _L, _D, _D, _P]),
    "ref_minidisk_initial_conserved": (None, [_P, _I, _L, _L, _P]),
    "ref_minidisk_mesh_create": (_P, [_P, _P, _I]),
    "ref_minidisk_mesh_rk2": (None, [_P, _D, _D, _I]),
    "ref_minidisk_mesh_get": (None, [_P, _P]),
    "ref_minidisk_mesh_destroy": (None, [_P]),
    "ref_mhd_recover_primitive": (_I, [_P, _P, _D, _P]),
    "ref_mhd_conserved_density": (None, [_P, _D, _P]),
    "ref_mhd_flux": (None, [_P, _I, _D, _P]),
    "ref_mhd_fastest_speed": (_D, [_P, _D]),
    "ref_mhd_riemann_hlle": (None, [_P, _P, _D, _I, _D, _P, _P]),
    "ref_mhd_advance_ct": (_I, [_I, _P, _P, _P, _P, _D, _D, _D, _P, _P, _P, _P, _P]),
    "ref_srhd_recover_primitive": (_I, [_P, _D, _D, _P]),
    "ref_srhd_conserved_density": (None, [_P, _D, _P]),
    "ref_srhd_riemann": (_I, [_P, _P, _I, _D, _I, _D, _P]),
}


def _load(path, sig):
    h = ctypes.CDLL(path)
    for name, (res, args) in sig.items():
        if hasattr(h, name):
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
    return h


_orc = None
_ref = None


def orc():
    global _orc
    if _orc is None:
        _orc = _load(ORACLE_PATH, _SIG)
    return _orc


def ref_available():
    return os.path.exists(REF_PATH)


def ref():
    global _ref
    if _ref is None:
        _ref = _load(REF_PATH, _REF_SIG)
    return _ref


def P(a):
    """ctypes pointer to a C-contiguous float64 numpy array."""
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)   # keeps the array alive for the call
