"""Device-clocked minidisk steps (m3p_minidisk_clock, m3p_minidisk_stage_dev):
the clock kernel against the reference's time step rule and the oracle's
binary orbit, device-clocked steps against host-timed ones, and hipGraph
replays of device-clocked steps against the same steps issued eagerly."""
import ctypes

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]

from oracle import P, orc  # noqa: E402


def rel_err(got, ref):
    """max over zones of |got - ref| / max|ref| per component (component axis = 1)."""
    scale = np.abs(ref).max(axis=tuple(i for i in range(ref.ndim) if i != 1), keepdims=True)
    return float((np.abs(got - ref) / np.maximum(scale, 1e-300)).max())


def _solver(precision="fp64", **kw):
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    args = dict(block_size=64, depth=2, domain_radius=2.0, softening_length=0.05, sink_radius=0.05)
    args.update(kw)
    data = MinidiskData.defaults(**args)
    _, u0 = problems.minidisk(data, seed=11)
    s = MinidiskSolver(data, precision)
    s.set_state(u0)
    return data, u0, s


@pytest.mark.parametrize("e, omega", [(0.0, 0.0), (0.0, 1.0), (0.4, 1.0), (0.8, -0.5)])
def test_clock_matches_reference_rule_and_oracle_orbit(e, omega):
    """dt and the next time equal the host formulas bit for bit (IEEE +, *, /
    and a correctly rounded sqrt taken on the host); the binary's state at t
    and t + dt agrees with the oracle's two-body solve (pinned against the
    reference's golden vectors) to 1e-15: only the device's sin and cos may
    differ from the host's, by an ulp."""
    from mara3p_amd import problems
    data, _, s = _solver(eccentricity=e, omega_frame=omega, mass_ratio=0.5)
    for t in (0.0, 0.013, 0.0495, 0.5, 3.7, 41.25):
        s.set_clock(t)
        from mara3p_amd._lib import check, lib
        check(lib().m3p_minidisk_clock(ctypes.byref(data), ctypes.c_void_p(s.clock.data_ptr()),
                                       ctypes.c_void_p(s.orbit.data_ptr()), 2,
                                       ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)))
        t_next, dt = s.clock_time()
        assert dt == problems.minidisk_dt(data, t)
        assert t_next == t * 0.5 + (t + dt) * (1.0 - 0.5)
        orbit = s.orbit.cpu().numpy().reshape(2, 6)
        for k, tk in enumerate((t, t + dt)):
            st = np.zeros(10)
            orc().orc_two_body_state(1.0, 1.0, data.mass_ratio, data.eccentricity, tk, data.omega_frame, P(st))
            ref = st[[0, 1, 2, 5, 6, 7]]
            assert np.abs(orbit[k] - ref).max() <= 1e-15, (t, k, orbit[k] - ref)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_clocked_rk2_steps_match_host_timed_steps(precision):
    """Five RK2 steps on the device clock against five host-timed steps (the
    host's dt rule and orbit): the same time sequence bit for bit, and states
    that differ only through the last bit of the orbit (<= 1e-13 relative)."""
    from mara3p_amd import problems
    data, u0, a = _solver(precision)
    _, _, b = _solver(precision)
    b.set_clock(0.0)
    t = 0.0
    for _ in range(5):
        a.step(problems.minidisk_dt(data, a.time))
        b.device_step(b.u, b.u2)
        b.u, b.u2 = b.u2, b.u
    a.check_errors()
    b.check_errors()
    tb, _ = b.clock_time()
    assert tb == a.time
    err = rel_err(b.state(), a.state())
    assert err <= 1e-13, err


def test_device_clocked_forward_euler_step():
    """rk_order 1: one stage per step (the reference driver's substep,
    problems/minidisk.cpp:172-186) and t + dt; equal to a host-timed first
    stage to 1e-13."""
    from mara3p_amd import problems
    data, u0, a = _solver()
    a.set_clock(0.25)
    a.device_step(a.u, a.u2, rk_order=1)
    t1, dt = a.clock_time()
    assert dt == problems.minidisk_dt(data, 0.25)
    assert t1 == 0.25 + dt
    _, _, b = _solver()
    b.stage(b.u, None, b.u2, 0.25, dt)
    a.check_errors()
    b.check_errors()
    err = rel_err(a.u2.cpu().numpy(), b.u2.cpu().numpy())
    assert err <= 1e-13, err


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
@pytest.mark.parametrize("rk_order", [2, 1])
def test_graph_replays_equal_eager_device_steps(precision, rk_order):
    """A hipGraph of two device-clocked steps, replayed three times with
    eager work and synchronisations between the replays, against the same
    six steps issued eagerly: every replay's state and clock bit for bit."""
    from mara3p_amd.solvers import Graph
    data, u0, g_s = _solver(precision)
    _, _, e_s = _solver(precision)
    g_s.set_clock(0.01)
    e_s.set_clock(0.01)
    stream = torch.cuda.Stream()
    with torch.cuda.stream(stream):
        with Graph() as g:
            g_s.device_step_pair(rk_order)
    torch.cuda.synchronize()
    # capture issues no work: the state and clock are still the initial ones
    assert np.array_equal(g_s.state(), e_s.state())
    for rep in range(3):
        with torch.cuda.stream(stream):
            g.replay()
        stream.synchronize()
        e_s.device_step_pair(rk_order)
        torch.cuda.synchronize()
        _ = torch.ones(1024, device="cuda").sum().item()      # eager work between replays
        assert np.array_equal(g_s.state(), e_s.state()), rep
        assert g_s.clock_time() == e_s.clock_time(), rep
    g_s.check_errors()
    e_s.check_errors()
    del g
