"""Error behaviour of the C ABI: the reference's exception cases map to status
codes (negative density -> std::invalid_argument in the reference,
src/physics_euler.hpp:158-161; shape errors, src/core_ndarray.hpp:714-732),
and every entry point rejects unusable arguments before launching anything."""
import ctypes

import numpy as np
import pytest

from mara3p_amd import _lib
from mara3p_amd._lib import (M3P_ERR_BAD_ARGUMENT, M3P_ERR_BAD_SHAPE, M3P_FP64, M3P_KIND_EULER2D, M3P_MIXED,
                             M3PError, NegativeDensityError, lib)


def test_tile_table_rejects_unknown_shapes_and_kinds():
    """No device needed: the tile table is host state."""
    assert lib().m3p_set_stage2d_tile(M3P_KIND_EULER2D, M3P_FP64, 48, 16) == M3P_ERR_BAD_ARGUMENT
    assert "not a compiled variant" in _lib.last_error()
    assert lib().m3p_set_stage2d_tile(7, M3P_FP64, 256, 16) == M3P_ERR_BAD_ARGUMENT
    tj, h = ctypes.c_int(0), ctypes.c_int(0)
    assert lib().m3p_get_stage2d_tile(M3P_KIND_EULER2D, M3P_MIXED, ctypes.byref(tj), ctypes.byref(h)) == 0
    assert (tj.value, h.value) == (256, 16)


def test_check_maps_status_to_reference_exceptions():
    with pytest.raises(NegativeDensityError):
        _lib.check(_lib.M3P_ERR_NEGATIVE_DENSITY)
    with pytest.raises(ValueError):        # std::invalid_argument in the reference
        _lib.check(_lib.M3P_ERR_NEGATIVE_DENSITY)
    with pytest.raises(M3PError):
        _lib.check(M3P_ERR_BAD_SHAPE)
    _lib.check(_lib.M3P_OK)


torch = pytest.importorskip("torch")
gpu = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")
def test_stage_argument_errors():
    f64 = dict(dtype=torch.float64, device="cuda")
    n = 64
    u = torch.ones((4, 5, n, n), **f64)
    out = torch.empty_like(u)
    nbr = torch.zeros((4, 4), dtype=torch.int32, device="cuda")
    L = lib()
    # output aliasing the input
    assert L.m3p_euler2d_stage(_p(u), _p(None), _p(u), _p(nbr), 4, n, 1e-3, _p(None), 1 / n, 0.5, 1.5, 1.4, 0,
                               M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    # no blocks
    assert L.m3p_euler2d_stage(_p(u), _p(None), _p(out), _p(nbr), 0, n, 1e-3, _p(None), 1 / n, 0.5, 1.5, 1.4, 0,
                               M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    # a block size no tile shape fits
    assert L.m3p_euler2d_stage(_p(u), _p(None), _p(out), _p(nbr), 1, 24, 1e-3, _p(None), 1 / 24, 0.5, 1.5, 1.4, 0,
                               M3P_FP64, _s()) == M3P_ERR_BAD_SHAPE
    # 1D: block size not supported
    u1 = torch.ones((2, 5, 48), **f64)
    o1 = torch.empty_like(u1)
    nb1 = torch.tensor([[-1, 1], [0, -1]], dtype=torch.int32, device="cuda")
    assert L.m3p_euler1d_stage(_p(u1), _p(None), _p(o1), _p(nb1), 2, 48, 1e-3, 1 / 96, 0.5, 1.5, 1.4,
                               M3P_FP64, _s()) == M3P_ERR_BAD_SHAPE
    # quad-tree leaves must be even and >= 16
    rec = torch.zeros((1, 4, 8), dtype=torch.int32, device="cuda")
    dl = torch.ones(1, **f64)
    u8 = torch.ones((1, 5, 8, 8), **f64)
    o8 = torch.empty_like(u8)
    assert L.m3p_amr_euler2d_stage(_p(u8), _p(None), _p(o8), _p(None), _p(rec), _p(dl), 1, 8, 1e-3, _p(None), 0.5,
                                   1.5, 1.4, 0, M3P_FP64, _s()) == M3P_ERR_BAD_SHAPE
    # negative viscosity
    data = _lib.MinidiskData.defaults(block_size=16, kinematic_viscosity=-1.0)
    uc = torch.ones((3, 16, 16), **f64)
    pe = torch.ones((3, 20, 20), **f64)
    oc = torch.empty_like(uc)
    assert L.m3p_minidisk_updated_conserved(_p(uc), _p(pe), _p(oc), 1, 0, 0, 0.0, 1e-3, ctypes.byref(data),
                                            M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    # fused CFL stages need cfl > 0 and an output for the time step
    dtn = torch.zeros(1, **f64)
    assert L.m3p_euler2d_stage_cfl(_p(u), _p(None), _p(out), _p(nbr), 4, n, 1e-3, _p(None), 1 / n, 0.5, 1.5, 1.4, 0,
                                   0.0, _p(dtn), M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    assert L.m3p_euler2d_stage_cfl(_p(u), _p(None), _p(out), _p(nbr), 4, n, 1e-3, _p(None), 1 / n, 0.5, 1.5, 1.4, 0,
                                   0.4, _p(None), M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    assert L.m3p_euler2d_stage_cfl(_p(u), _p(None), _p(u), _p(nbr), 4, n, 1e-3, _p(None), 1 / n, 0.5, 1.5, 1.4, 0,
                                   0.4, _p(dtn), M3P_FP64, _s()) == M3P_ERR_BAD_ARGUMENT
    # the library still works after rejected calls
    torch.cuda.synchronize()
    code = ctypes.c_int(0)
    assert L.m3p_check_errors(_s(), ctypes.byref(code), None, None) == 0 and code.value == 0


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_negative_density_in_ensemble_and_tree(precision):
    """A negative density anywhere in an ensemble or a quad-tree stage latches
    the reference's error (block index reported), and the latch clears."""
    from mara3p_amd import amr
    from mara3p_amd.mesh import EnsembleMesh1D
    from mara3p_amd.problems import riemann_ensemble, sedov_amr
    from mara3p_amd.solvers import AmrEuler2DSolver, Euler1DEnsembleSolver
    m, u = riemann_ensemble(count=8, block_size=64, lmin=128, lmax=512, seed=2)
    u = u.copy()
    u[3, 0, 17] = -0.5
    s = Euler1DEnsembleSolver(m, precision=precision)
    s.set_state(u)
    s.step(1e-5)
    with pytest.raises(NegativeDensityError) as e:
        s.check_errors()
    assert "block" in str(e.value)
    s.check_errors()
    t = amr.nested_tree(3, 1.5)
    ua = sedov_amr(t, 16).copy()
    ua[5, 0, 3, 4] = -1.0
    a = AmrEuler2DSolver(t, 16, precision=precision)
    a.set_state(ua)
    a.step(1e-6)
    with pytest.raises(NegativeDensityError):
        a.check_errors()
    a.check_errors()


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")
def test_single_block_and_smallest_meshes():
    """Edge sizes: a 1 x 1 periodic mesh of 16^2 (every neighbour is the block
    itself) equals the oracle; a one-block 1D chain takes only the lower
    outflow (the reference's quirk) and equals the oracle."""
    from oracle import P, orc
    from mara3p_amd import problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import Euler1DSolver, Euler2DSolver
    mesh, u0 = problems.kelvin_helmholtz(depth=0, block_size=16, seed=4)
    s = Euler2DSolver(0, 16)
    s.set_state(u0)
    dt = 0.2 * s.dl
    s.step(dt)
    ref = to_reference_layout(u0).copy()
    assert orc().orc_euler2d_mesh_rk2(P(ref), 1, 16, dt, s.dl, 1.5, 5.0 / 3.0, 0, 1, 1) == 0
    assert np.array_equal(s.state(), to_device_layout(ref))
    u1 = problems.sod(nblk=1, block_size=64)
    c = Euler1DSolver(1, 64)
    c.set_state(u1)
    c.step(1e-3)
    r1 = np.ascontiguousarray(np.moveaxis(u1, 1, -1))
    assert orc().orc_euler1d_mesh_rk2(P(r1), 1, 64, 1e-3, c.dl, 1.5, 1.4, 1, 1) == 0
    assert np.array_equal(c.state(), np.moveaxis(r1, -1, 1))
