"""The mhd_scheme_v2 array functions (src/scheme_mhd_v2.hpp:71-210) on
device-resident torch tensors, over the C ABI (m3p_mhd_v2_*).

Names and argument meaning follow the reference: a "face" argument is the
3-tuple (bf1, bf2, bf3), Godunov data is the 3-tuple (gf1, gf2, gf3) and edge
fields the 3-tuple (ee1, ee2, ee3).  Arrays are float64 on the device,
components outermost: conserved (5, n1, n2, n3), primitive (8, n1, n2, n3),
Godunov data (8, face shape); faces and edges are scalar arrays in the
reference's shapes (x faces (n1+1, n2, n3), x edges (n1, n2+1, n3+1), ...).
The adiabatic index is 5/3, as in the reference (src/scheme_mhd_v2.cpp:48).
Each function returns new tensors, as the reference returns new arrays.
"""
import ctypes

from . import _lib

GAMMA_LAW_INDEX = 5.0 / 3.0


def _torch():
    import torch
    return torch


def _p(t):
    return ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _check_array(t, shape, what):
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()):
        raise TypeError("%s: expected a contiguous float64 device tensor" % what)
    if tuple(t.shape) != tuple(shape):
        raise ValueError("%s: shape %s, expected %s" % (what, tuple(t.shape), tuple(shape)))


def _empty(*shape):
    torch = _torch()
    return torch.empty(shape, dtype=torch.float64, device="cuda")


def face_shapes(n1, n2, n3):
    """Shapes of the three face arrays of cells (n1, n2, n3)."""
    return (n1 + 1, n2, n3), (n1, n2 + 1, n3), (n1, n2, n3 + 1)


def edge_shapes(n1, n2, n3):
    """Shapes of the three edge arrays of cells (n1, n2, n3)."""
    return (n1, n2 + 1, n3 + 1), (n1 + 1, n2, n3 + 1), (n1 + 1, n2 + 1, n3)


def _check_faces(bf, n, what):
    for a, (t, s) in enumerate(zip(bf, face_shapes(*n))):
        _check_array(t, s, "%s: bf%d" % (what, a + 1))


def primitive_array(uc, bf):
    """primitive_array (src/scheme_mhd_v2.cpp:75-81): raises ValueError where
    mhd::recover_primitive throws (negative density or pressure)."""
    n = tuple(uc.shape[1:])
    _check_array(uc, (5,) + n, "primitive_array: uc")
    _check_faces(bf, n, "primitive_array")
    pc = _empty(8, *n)
    _lib.check(_lib.lib().m3p_mhd_v2_primitive_array(_p(uc), _p(bf[0]), _p(bf[1]), _p(bf[2]), *n, _p(pc), _stream()))
    return pc


def godunov_fluxes(pc, bf):
    """godunov_fluxes (:168-185): pc extended by one zone on every axis, bf by
    one zone on the transverse axes; returns (gf1, gf2, gf3)."""
    m = tuple(pc.shape[1:])
    _check_array(pc, (8,) + m, "godunov_fluxes: pc")
    shapes = [tuple(s - (a == b) for b, s in enumerate(m)) for a in range(3)]
    for a in range(3):
        _check_array(bf[a], shapes[a], "godunov_fluxes: bf%d" % (a + 1))
    gf = tuple(_empty(8, *s) for s in shapes)
    _lib.check(_lib.lib().m3p_mhd_v2_godunov_fluxes(_p(pc), _p(bf[0]), _p(bf[1]), _p(bf[2]), *m,
                                                    _p(gf[0]), _p(gf[1]), _p(gf[2]), _stream()))
    return gf


def electromotive_forces(gf):
    """electromotive_forces (:191-214): returns (ee1, ee2, ee3)."""
    m = (gf[1].shape[1], gf[0].shape[2], gf[0].shape[3])
    for a in range(3):
        _check_array(gf[a], (8,) + tuple(s - (a == b) for b, s in enumerate(m)), "electromotive_forces: gf%d" % (a + 1))
    shapes = [tuple(s - (a != b) for b, s in enumerate(m)) for a in range(3)]
    ee = tuple(_empty(*s) for s in shapes)
    _lib.check(_lib.lib().m3p_mhd_v2_electromotive_forces(_p(gf[0]), _p(gf[1]), _p(gf[2]), *m,
                                                          _p(ee[0]), _p(ee[1]), _p(ee[2]), _stream()))
    return ee


def updated_conserved_density(uc, gf, dt, dl):
    """updated_conserved_density (:220-241): gf extended by one zone on the
    transverse axes, as godunov_fluxes returns it for pc extended by one."""
    n = tuple(uc.shape[1:])
    _check_array(uc, (5,) + n, "updated_conserved_density: uc")
    for a in range(3):
        s = tuple(x + 1 if a == b else x + 2 for b, x in enumerate(n))
        _check_array(gf[a], (8,) + s, "updated_conserved_density: gf%d" % (a + 1))
    out = _empty(5, *n)
    _lib.check(_lib.lib().m3p_mhd_v2_updated_conserved_density(_p(uc), _p(gf[0]), _p(gf[1]), _p(gf[2]), *n,
                                                               float(dt), float(dl), _p(out), _stream()))
    return out


def curl(ee, dl):
    """curl (:247-258): faces from edges (n1, n2+1, n3+1), ... over a unit time."""
    n = (ee[0].shape[0], ee[1].shape[1], ee[2].shape[2])
    for a, s in enumerate(edge_shapes(*n)):
        _check_array(ee[a], s, "curl: ee%d" % (a + 1))
    bf = tuple(_empty(*s) for s in face_shapes(*n))
    _lib.check(_lib.lib().m3p_mhd_v2_curl(_p(ee[0]), _p(ee[1]), _p(ee[2]), *n, float(dl),
                                          _p(bf[0]), _p(bf[1]), _p(bf[2]), _stream()))
    return bf


def updated_magnetic_flux_density(bf, ee, dt, dl):
    """updated_magnetic_flux_density (:264-278): ee as electromotive_forces
    returns it for Godunov data extended by one zone."""
    n = (bf[1].shape[0], bf[0].shape[1], bf[0].shape[2])
    _check_faces(bf, n, "updated_magnetic_flux_density")
    for a, s in enumerate(edge_shapes(*n)):
        _check_array(ee[a], tuple(x + 2 if a == b else x for b, x in enumerate(s)),
                     "updated_magnetic_flux_density: ee%d" % (a + 1))
    out = tuple(_empty(*s) for s in face_shapes(*n))
    _lib.check(_lib.lib().m3p_mhd_v2_updated_magnetic_flux_density(
        _p(bf[0]), _p(bf[1]), _p(bf[2]), _p(ee[0]), _p(ee[1]), _p(ee[2]), *n, float(dt), float(dl),
        _p(out[0]), _p(out[1]), _p(out[2]), _stream()))
    return out


def _pointer_list(tensors):
    return (ctypes.c_void_p * 27)(*[t.data_ptr() for t in tensors])


def extend_cell_primitive_variables(pc_vector, block_size, count):
    """extend_cell_primitive_variables (:126-135): pc_vector is the 27 blocks
    of a 3 x 3 x 3 neighbourhood in the order of neighbors_27."""
    if len(pc_vector) != 27:
        raise ValueError("extend_cell_primitive_variables: need 27 blocks")
    for t in pc_vector:
        _check_array(t, (8, block_size, block_size, block_size), "extend_cell_primitive_variables: block")
    m = block_size + 2 * count
    out = _empty(8, m, m, m)
    _lib.check(_lib.lib().m3p_mhd_v2_extend_cell_primitive_variables(_pointer_list(pc_vector), block_size, count,
                                                                     _p(out), _stream()))
    return out


def extend_face_magnetic_flux_density(bf_vector, block_size, count):
    """extend_face_magnetic_flux_density (:141-162): bf_vector is 27 face
    triplets, three runs of 9: run a holds the 3 x 3 square of blocks normal to
    axis a in the order of neighbors_9(..., a), and only its axis-a faces are
    read (as the reference does)."""
    if len(bf_vector) != 27:
        raise ValueError("extend_face_magnetic_flux_density: need 27 face triplets")
    picked = [bf_vector[9 * a + m][a] for a in range(3) for m in range(9)]
    shapes = face_shapes(block_size, block_size, block_size)
    for a in range(3):
        for t in picked[9 * a:9 * a + 9]:
            _check_array(t, shapes[a], "extend_face_magnetic_flux_density: bf%d" % (a + 1))
    m = block_size + 2 * count
    out = tuple(_empty(*[block_size + 1 if a == b else m for b in range(3)]) for a in range(3))
    _lib.check(_lib.lib().m3p_mhd_v2_extend_face_magnetic_flux_density(
        _pointer_list(picked), block_size, count, _p(out[0]), _p(out[1]), _p(out[2]), _stream()))
    return out


def neighbors_27(base_index, extent):
    """mesh::neighbors_27 (src/mesh_cartesian_3d.hpp:355-364): the 27 block
    indices around base_index, row-major over the offsets, wrapped periodically."""
    return [tuple((b + d + e - 1) % e for b, d, e in zip(base_index, (di, dj, dk), extent))
            for di in range(3) for dj in range(3) for dk in range(3)]


def neighbors_9(base_index, extent, axis):
    """mesh::neighbors_9 (src/mesh_cartesian_3d.hpp:384-395): the 3 x 3 square
    of block indices normal to axis (0, 1, 2) through base_index, row-major."""
    space = [1 if a == axis else 3 for a in range(3)]
    offset = [0 if a == axis else 1 for a in range(3)]
    return [tuple((b + d + e - o) % e for b, d, e, o in zip(base_index, (di, dj, dk), extent, offset))
            for di in range(space[0]) for dj in range(space[1]) for dk in range(space[2])]


def advance_block_mesh(uc_blocks, bf_blocks, extent, dt, dl):
    """One first-order CT update of a periodic mesh of blocks through the v2
    functions: for each block, extend its primitives and faces by one zone
    from its neighbours, then Godunov fluxes, edge fields and the conserved
    and face updates (the composition src/scheme_mhd_v1.cpp:265-299 performs
    with plm_theta = 0).  uc_blocks / bf_blocks: dicts keyed by block index
    (i, j, k) in [0, extent); returns the updated dicts."""
    pc = {b: primitive_array(uc_blocks[b], bf_blocks[b]) for b in uc_blocks}
    bs = next(iter(pc.values())).shape[1]
    uo, bo = {}, {}
    for b in uc_blocks:
        pce = extend_cell_primitive_variables([pc[x] for x in neighbors_27(b, extent)], bs, 1)
        bfe = extend_face_magnetic_flux_density(
            [bf_blocks[x] for a in range(3) for x in neighbors_9(b, extent, a)], bs, 1)
        gf = godunov_fluxes(pce, bfe)
        ee = electromotive_forces(gf)
        uo[b] = updated_conserved_density(uc_blocks[b], gf, dt, dl)
        bo[b] = updated_magnetic_flux_density(bf_blocks[b], ee, dt, dl)
    return uo, bo
