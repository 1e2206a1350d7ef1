"""The fused euler2d RK2 step (m3p_euler2d_rk2_step, one launch per step, the
first stage's state never stored) against the two stage launches it
replaces: the same arithmetic on the same operands, so every value must be
equal (==), with HLLE and HLLC, on meshes of every supported block size, on
a single block that is its own neighbour and on the full 4096^2
Kelvin-Helmholtz mesh.  The form with the CFL reduction fused in leaves the
time step that the separate reduction computes from its output, bit for bit,
and its hipGraph replays equal eager two-stage steps with separate
reductions."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _pair(depth, n, hllc=True, seed=3):
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=depth, block_size=n, seed=seed)
    a = Euler2DSolver(depth, n, gamma=5.0 / 3.0, plm_theta=1.5, precision="mixed", hllc=hllc)
    b = Euler2DSolver(depth, n, gamma=5.0 / 3.0, plm_theta=1.5, precision="mixed", hllc=hllc)
    a.set_state(u0)
    b.set_state(u0)
    return a, b


@pytest.mark.parametrize("hllc", [True, False])
@pytest.mark.parametrize("depth, n", [(0, 64), (2, 64), (1, 128), (1, 256), (3, 64)])
def test_fused_rk2_equals_two_stages(depth, n, hllc):
    """Three RK2 steps at the CFL time step of the initial state, fused
    against two stage launches each: equal in every value; depth 0 is one
    block that is its own neighbour on both axes."""
    a, b = _pair(depth, n, hllc, seed=5 + depth)
    dt = float(a.cfl_dt(0.4).item())
    for _ in range(3):
        a.step(dt)
        b.step_fused(dt)
    a.check_errors()
    b.check_errors()
    ua, ub = a.state(), b.state()
    assert np.array_equal(ua, ub), float(np.abs(ua - ub).max())


def test_fused_rk2_equals_two_stages_full_size():
    """The Kelvin-Helmholtz configuration (16 x 16 blocks of 256^2, HLLC):
    two steps, equal in every value."""
    a, b = _pair(4, 256, True, seed=1)
    dt = float(a.cfl_dt(0.4).item())
    for _ in range(2):
        a.step(dt)
        b.step_fused(dt)
    a.check_errors()
    b.check_errors()
    assert np.array_equal(a.state(), b.state())


def test_fused_rk2_cfl_output_and_dt():
    """The _cfl form: the same output as the plain fused step, and dt_next
    equal to the separate reduction of that output (m3p_euler_cfl_dt)."""
    a, _ = _pair(2, 64, True, seed=7)
    dt = a.cfl_dt(0.4)
    out_plain = torch.empty_like(a.u)
    out_cfl = torch.empty_like(a.u)
    dt_next = torch.zeros(1, dtype=torch.float64, device="cuda")
    a.rk2_step(a.u, out_plain, 0.0, dt_dev=dt)
    a.rk2_step(a.u, out_cfl, 0.0, dt_dev=dt, cfl=0.4, dt_next=dt_next)
    a.check_errors()
    assert torch.equal(out_plain, out_cfl)
    a.u = out_cfl
    sep = float(a.cfl_dt(0.4).item())
    assert float(dt_next.item()) == sep
    assert 0.0 < sep < 1.0


def test_fused_rk2_cfl_graph_replays_match_eager_separate():
    """A hipGraph of two fused CFL steps (each launch reduces the next
    step's time step into the solver's dt), replayed three times with eager
    work between the replays, against the same steps issued eagerly as two
    stage launches with a separate CFL reduction before each step: the state
    and the time step after every replay, bit for bit."""
    from mara3p_amd.solvers import Graph, cfl_step_pair, cfl_step_pair_rk2
    g_s, e_s = _pair(2, 64, True, seed=11)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        g_s.cfl_dt(0.4)
        with Graph() as g:
            cfl_step_pair_rk2(g_s, 0.4)
    stream.synchronize()
    torch.cuda.synchronize()
    for rep in range(3):
        with torch.cuda.stream(stream):
            g.replay()
        stream.synchronize()
        cfl_step_pair(e_s, 0.4)
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()      # eager work between replays
        assert np.array_equal(g_s.state(), e_s.state()), rep
        assert float(g_s.dt_dev.item()) == float(e_s.cfl_dt(0.4).item()), rep
    g_s.check_errors()
    e_s.check_errors()
    del g


def test_fused_rk2_rejects_what_it_does_not_compile():
    """FP64, blocks other than 64/128/256 zones a side and aliasing are
    refused with an error, before any launch."""
    from mara3p_amd._lib import lib
    from mara3p_amd.solvers import Euler2DSolver
    L = lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for n, prec in ((64, "fp64"), (32, "mixed")):
        s = Euler2DSolver(1, n, precision=prec, hllc=True)
        st = L.m3p_euler2d_rk2_step(ctypes.c_void_p(s.u.data_ptr()), ctypes.c_void_p(s.u2.data_ptr()),
                                    ctypes.c_void_p(s.nbr.data_ptr()), s.mesh.nblk, n, 1e-3, ctypes.c_void_p(0),
                                    s.dl, 0.5, s.theta, s.gamma, 1, s.precision, stream)
        assert st != 0, (n, prec)
    s = Euler2DSolver(1, 64, precision="mixed", hllc=True)
    st = L.m3p_euler2d_rk2_step(ctypes.c_void_p(s.u.data_ptr()), ctypes.c_void_p(s.u.data_ptr()),
                                ctypes.c_void_p(s.nbr.data_ptr()), s.mesh.nblk, 64, 1e-3, ctypes.c_void_p(0),
                                s.dl, 0.5, s.theta, s.gamma, 1, s.precision, stream)
    assert st != 0
