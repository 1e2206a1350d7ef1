"""Stages with the CFL reduction fused in (m3p_euler2d_stage_cfl,
m3p_amr_euler2d_stage_cfl, m3p_euler1d_stage_blocks_cfl): the stage output
equals the plain stage's bit for bit, and the time step they leave equals
the separate reduction (m3p_euler_cfl_dt / m3p_euler_cfl_dt_blocks) of that
output bit for bit; captured in a hipGraph, fused CFL steps replay the eager
steps with separate reductions exactly."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def _solvers(kind, precision):
    from mara3p_amd import amr, problems
    from mara3p_amd.solvers import AmrEuler2DSolver, Euler1DEnsembleSolver, Euler2DSolver
    if kind == "uniform":
        _, u0 = problems.kelvin_helmholtz(depth=1, block_size=64, seed=3)
        make = lambda: Euler2DSolver(1, 64, precision=precision, hllc=True)
    elif kind == "tree":
        t = amr.nested_tree(4, 1.5)
        u0 = problems.sedov_amr(t, 16)
        make = lambda: AmrEuler2DSolver(t, 16, precision=precision)
    else:
        mesh, u0 = problems.riemann_ensemble(count=12, block_size=64, lmin=128, lmax=512, seed=4)
        make = lambda: Euler1DEnsembleSolver(mesh, precision=precision)
    a, b = make(), make()
    a.set_state(u0)
    b.set_state(u0)
    return a, b


@pytest.mark.parametrize("kind", ["uniform", "tree", "ensemble"])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_stage_output_and_dt_equal_separate(kind, precision):
    s, _ = _solvers(kind, precision)
    dt = s.cfl_dt(0.4)
    out_plain = torch.empty_like(s.u)
    out_fused = torch.empty_like(s.u)
    dt_next = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.stage(s.u, None, s.u1, 0.0, dt_dev=dt)
    s.stage(s.u1, s.u, out_plain, 0.0, dt_dev=dt)
    s.stage(s.u1, s.u, out_fused, 0.0, dt_dev=dt, cfl=0.4, dt_next=dt_next)
    # the first stage alone, fused, as well
    u1_fused = torch.empty_like(s.u)
    dt1 = torch.zeros(1, dtype=torch.float64, device="cuda")
    s.stage(s.u, None, u1_fused, 0.0, dt_dev=dt, cfl=0.4, dt_next=dt1)
    s.check_errors()
    assert torch.equal(out_plain, out_fused)
    assert torch.equal(s.u1, u1_fused)
    s.u = out_plain
    sep = float(s.cfl_dt(0.4).item())
    assert float(dt_next.item()) == sep
    s.u = u1_fused
    assert float(dt1.item()) == float(s.cfl_dt(0.4).item())
    assert 0.0 < sep < 1.0


@pytest.mark.parametrize("kind", ["uniform", "tree", "ensemble"])
def test_fused_cfl_graph_replays_match_eager_separate(kind):
    """Three replays of a captured pair of fused CFL steps (eager work and
    synchronisations between them) equal six eager steps with a separate CFL
    reduction before each."""
    from mara3p_amd.solvers import Graph, cfl_step_pair, cfl_step_pair_fused
    a, b = _solvers(kind, "mixed")
    for _ in range(3):
        cfl_step_pair(a, 0.4)
    a.check_errors()
    b.cfl_dt(0.4)
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    with torch.cuda.stream(stream):
        with Graph() as g:
            cfl_step_pair_fused(b, 0.4)
        for _ in range(3):
            g.replay()
            stream.synchronize()
            scratch = torch.ones(1000, device="cuda").sum()    # eager work between replays
            assert float(scratch.item()) == 1000.0
    torch.cuda.synchronize()
    b.check_errors()
    assert np.array_equal(a.state(), b.state())
    # both leave the time step of the final state
    ref = float(a.cfl_dt(0.4).item())
    assert float(b.dt_dev.item()) == ref
