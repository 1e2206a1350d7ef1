"""The fused RK2 step (m3p_minidisk_rk2_step, one launch per step, the first
stage's state never stored) against the two stage launches it replaces: the
same arithmetic on the same operands, so every value must be equal (==), on
meshes of every supported block size, on the full-size minidisk mesh, on a
single block that is its own neighbour, on the device clock and under
hipGraph replay."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _pair(n, depth, seed=3, **kw):
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    args = dict(block_size=n, depth=depth, domain_radius=2.0, softening_length=0.05, sink_radius=0.05)
    args.update(kw)
    data = MinidiskData.defaults(**args)
    _, u0 = problems.minidisk(data, seed=seed)
    a, b = MinidiskSolver(data, "mixed"), MinidiskSolver(data, "mixed")
    a.set_state(u0)
    b.set_state(u0)
    return data, u0, a, b


@pytest.mark.parametrize("n, depth", [(64, 0), (64, 2), (128, 1), (256, 1), (64, 3)])
def test_fused_rk2_equals_two_stages(n, depth):
    """Three RK2 steps, fused against two stage launches each, from the same
    state: equal in every value; depth 0 is one block that is its own
    neighbour on both axes."""
    from mara3p_amd import problems
    data, _, a, b = _pair(n, depth, seed=5 + depth)
    for _ in range(3):
        dt = problems.minidisk_dt(data, a.time)
        a.step(dt)
        b.step_fused(dt)
    a.check_errors()
    b.check_errors()
    assert a.time == b.time
    ua, ub = a.state(), b.state()
    assert np.array_equal(ua, ub), float(np.abs(ua - ub).max())


def test_fused_rk2_equals_two_stages_full_size():
    """The full minidisk configuration (16 x 16 blocks of 256^2, [-8, 8]^2,
    softening 0.05, Mach 10, sink radius 0.05): two steps, equal in every
    value."""
    from mara3p_amd import problems
    data, _, a, b = _pair(256, 4, seed=1, domain_radius=8.0)
    for _ in range(2):
        dt = problems.minidisk_dt(data, a.time)
        a.step(dt)
        b.step_fused(dt)
    a.check_errors()
    b.check_errors()
    assert np.array_equal(a.state(), b.state())


def test_fused_rk2_rotating_eccentric_binary():
    """An eccentric binary of unequal masses in a rotating frame, with the
    sink and buffer terms active at the domain edge (domain radius 1.5 with
    a 0.2 buffer): equal in every value after two steps."""
    from mara3p_amd import problems
    data, _, a, b = _pair(64, 2, seed=9, domain_radius=1.5, eccentricity=0.3, mass_ratio=0.6, omega_frame=0.7,
                          softening_length=0.02, sink_radius=0.05, sink_rate=10.0, buffer_rate=10.0,
                          buffer_scale=0.2)
    for _ in range(2):
        dt = problems.minidisk_dt(data, a.time)
        a.step(dt)
        b.step_fused(dt)
    assert np.array_equal(a.state(), b.state())


@pytest.mark.parametrize("fused_graph", [False, True])
def test_fused_device_clocked_steps(fused_graph):
    """Device-clocked fused steps against device-clocked stage launches
    (equal in every value and the same clock), eagerly and as three hipGraph
    replays of two steps."""
    from mara3p_amd.solvers import Graph
    _, _, a, b = _pair(64, 2, seed=12)
    a.set_clock(0.02)
    b.set_clock(0.02)
    if fused_graph:
        stream = torch.cuda.Stream()
        with torch.cuda.stream(stream):
            with Graph() as g:
                b.device_step_pair(2, fused=True)
        torch.cuda.synchronize()
    for rep in range(3):
        a.device_step_pair(2, fused=False)
        if fused_graph:
            with torch.cuda.stream(stream):
                g.replay()
            stream.synchronize()
        else:
            b.device_step_pair(2, fused=True)
        torch.cuda.synchronize()
        assert np.array_equal(a.state(), b.state()), rep
        assert a.clock_time() == b.clock_time(), rep
    a.check_errors()
    b.check_errors()


def test_fused_rk2_rejects_what_it_does_not_compile():
    """FP64, viscosity, block sizes other than 64/128/256 and aliasing are
    refused with an error, before any launch."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd._lib import lib
    from mara3p_amd.solvers import MinidiskSolver
    L = lib()
    for kw, prec in ((dict(block_size=64), "fp64"), (dict(block_size=64, kinematic_viscosity=0.01), "mixed"),
                     (dict(block_size=32), "mixed")):
        data = MinidiskData.defaults(depth=1, **kw)
        s = MinidiskSolver(data, prec)
        st = L.m3p_minidisk_rk2_step(ctypes.c_void_p(s.u.data_ptr()), ctypes.c_void_p(s.u2.data_ptr()),
                                     ctypes.c_void_p(s.nbr.data_ptr()), ctypes.c_void_p(s.bcoord.data_ptr()),
                                     s.mesh.nblk, 0.0, 1e-3, 0.5, ctypes.byref(s.data), s.precision,
                                     ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
        assert st != 0, (kw, prec)
    data = MinidiskData.defaults(depth=1, block_size=64)
    s = MinidiskSolver(data, "mixed")
    st = L.m3p_minidisk_rk2_step(ctypes.c_void_p(s.u.data_ptr()), ctypes.c_void_p(s.u.data_ptr()),
                                 ctypes.c_void_p(s.nbr.data_ptr()), ctypes.c_void_p(s.bcoord.data_ptr()),
                                 s.mesh.nblk, 0.0, 1e-3, 0.5, ctypes.byref(s.data), s.precision,
                                 ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert st != 0
