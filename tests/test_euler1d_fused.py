"""The fused 1D RK2 step (m3p_euler1d_rk2_step_blocks, one launch per step,
the first stage's state never stored) against the two work-list stage
launches it replaces: the same arithmetic on the same operands, so every
value must be equal (==), for every block size, on the full 512-problem
ensemble, on a ring (periodic) chain and at outflow ends.  The _cfl form
leaves the time step the separate reduction computes from its output, bit
for bit, and its hipGraph replays equal eager two-stage steps with separate
reductions.  Tables whose prev/next links are not mutual are refused."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _pair(count, n, lmin, lmax, seed=3):
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler1DEnsembleSolver
    mesh, u0 = problems.riemann_ensemble(count=count, block_size=n, lmin=lmin, lmax=lmax, seed=seed)
    a = Euler1DEnsembleSolver(mesh, gamma=1.4, plm_theta=1.5, precision="mixed")
    b = Euler1DEnsembleSolver(mesh, gamma=1.4, plm_theta=1.5, precision="mixed")
    a.set_state(u0)
    b.set_state(u0)
    return a, b


def _steps_equal(a, b, steps):
    dt = float(a.cfl_dt(0.4).item())
    for _ in range(steps):
        a.step(dt)
        b.step_fused(dt)
    a.check_errors()
    b.check_errors()
    ua, ub = a.state(), b.state()
    assert np.array_equal(ua, ub), float(np.abs(ua - ub).max())


@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512])
def test_fused_rk2_equals_two_stages(n):
    """Three RK2 steps, fused against two stage launches each, on an
    ensemble of 12 problems of 2 to 16 blocks (outflow ends, one cell
    spacing per problem): equal in every value.  n = 512 runs two tiles of
    256 zones per block."""
    a, b = _pair(12, n, 2 * n, 16 * n, seed=5 + n)
    _steps_equal(a, b, 3)


def test_fused_rk2_equals_two_stages_full_ensemble():
    """The 512-problem ensemble of 512 to 8192 zones in blocks of 64: two
    steps, equal in every value."""
    a, b = _pair(512, 64, 512, 8192, seed=1)
    _steps_equal(a, b, 2)


def test_fused_rk2_ring():
    """One problem of 6 blocks closed into a ring (periodic ends, the links
    rewritten on both solvers): equal in every value after three steps."""
    a, b = _pair(1, 64, 384, 384, seed=9)
    for s in (a, b):
        nbr = s.nbr.clone()
        nbr[0, 0] = s.mesh.nblk - 1
        nbr[s.mesh.nblk - 1, 1] = 0
        s.nbr = nbr
    _steps_equal(a, b, 3)


def test_fused_rk2_cfl_output_and_dt():
    """The _cfl form: the output of the plain fused step, and dt_next equal
    to the separate reduction of that output (m3p_euler_cfl_dt_blocks)."""
    a, _ = _pair(16, 64, 128, 1024, seed=7)
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
    """A hipGraph of two fused CFL steps replayed three times with eager work
    between the replays, against the same steps issued eagerly as two stage
    launches with a separate CFL reduction before each step: the state and
    the time step after every replay, bit for bit."""
    from mara3p_amd.solvers import Graph, cfl_step_pair, cfl_step_pair_rk2
    g_s, e_s = _pair(16, 64, 128, 1024, seed=11)
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


def test_fused_rk2_refuses_what_it_does_not_compute():
    """FP64, a block size outside the list and aliasing are refused before
    any launch; a one-block problem (prev outflow, next itself: links that
    are not mutual) is refused by the solver on the host and, through the C
    ABI, reported by the kernel as M3P_ERR_BAD_SHAPE."""
    from mara3p_amd import problems
    from mara3p_amd._lib import M3PError, lib
    from mara3p_amd.mesh import EnsembleMesh1D
    from mara3p_amd.solvers import Euler1DEnsembleSolver
    L = lib()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)

    def call(s, u, out, n):
        return L.m3p_euler1d_rk2_step_blocks(ctypes.c_void_p(u.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                             ctypes.c_void_p(s.nbr.data_ptr()), ctypes.c_void_p(s.dl.data_ptr()),
                                             s.mesh.nblk, n, 1e-4, ctypes.c_void_p(0), 0.5, s.theta, s.gamma,
                                             s.precision, stream)
    mesh, u0 = problems.riemann_ensemble(count=4, block_size=64, lmin=128, lmax=256, seed=2)
    s64 = Euler1DEnsembleSolver(mesh, precision="fp64")
    assert call(s64, s64.u, s64.u2, 64) != 0
    sm = Euler1DEnsembleSolver(mesh, precision="mixed")
    assert call(sm, sm.u, sm.u2, 48) != 0
    assert call(sm, sm.u, sm.u, 64) != 0
    # one problem of one block, one of three
    one = Euler1DEnsembleSolver(EnsembleMesh1D([64, 192], 64), precision="mixed")
    assert not one.links_mutual()
    with pytest.raises(ValueError):
        one.step_fused(1e-4)
    one.u.copy_(torch.as_tensor(np.concatenate([u0[:1], u0[:3]]), device="cuda"))
    assert call(one, one.u, one.u2, 64) == 0
    with pytest.raises(M3PError) as e:
        one.check_errors()
    assert e.value.code == 4
