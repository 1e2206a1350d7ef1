"""Viscous isothermal updates: the kinematic-viscosity stress of
godunov_and_viscous_fluxes (src/module_locally_isothermal.cpp:106-174,
problems/minidisk_scheme.cpp:220-287).  The oracle's restatement is pinned
bitwise against the compiled reference in tests/test_oracle.py (golden
fixture minidisk_visc_block.npz and the live-reference test); here the HIP
per-zone kernel is checked against the oracle."""
import ctypes

import numpy as np
import pytest

from oracle import OrcMinidisk, P, orc

torch = pytest.importorskip("torch")
pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def dev(a):
    return torch.as_tensor(c(a), device="cuda")


def ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def rel_err(got, ref):
    """max over zones of |got - ref| / max|ref| per component (component axis = 1)."""
    scale = np.abs(ref).max(axis=tuple(i for i in range(ref.ndim) if i != 1), keepdims=True)
    return float((np.abs(got - ref) / np.maximum(scale, 1e-300)).max())


def _random_block(n, rng):
    uc = c(np.abs(rng.normal(1, 0.1, (n, n, 3))))
    pe = c(np.concatenate([rng.uniform(0.5, 1.5, (n + 4, n + 4, 1)), rng.normal(0, 1, (n + 4, n + 4, 2))], -1))
    return uc, pe


@pytest.mark.parametrize("nu", [1e-3, 0.2])
@pytest.mark.parametrize("n", [16, 64])
def test_minidisk_block_viscous_fp64(nu, n):
    """Per-block minidisk::updated_conserved with nu != 0 (pe corners are read)
    against the oracle, 1e-12 relative (the sources call exp/pow/tanh)."""
    import mara3p_amd as m3p
    rng = np.random.default_rng(int(nu * 1000) + n)
    level, bi, bj = 2, 1, 3
    data = m3p.MinidiskData.defaults(block_size=n, depth=level, domain_radius=1.5, softening_length=0.05,
                                     sink_radius=0.05, kinematic_viscosity=nu)
    uc, pe = _random_block(n, rng)
    ref = np.zeros_like(uc)
    orc().orc_minidisk_updated_conserved(P(uc), P(pe), ctypes.byref(OrcMinidisk.from_data(data)), level, bi, bj,
                                         0.4, 1e-3, P(ref))
    out = torch.empty((3, n, n), dtype=torch.float64, device="cuda")
    d_uc, d_pe = dev(np.moveaxis(uc, -1, 0)), dev(np.moveaxis(pe, -1, 0))
    m3p.check(m3p.lib().m3p_minidisk_updated_conserved(ptr(d_uc), ptr(d_pe), ptr(out), level, bi, bj, 0.4, 1e-3,
                                                       ctypes.byref(data), m3p.M3P_FP64, stream()))
    got = np.moveaxis(out.cpu().numpy(), 0, -1)
    err = rel_err(got[None].transpose(0, 3, 1, 2), ref[None].transpose(0, 3, 1, 2))
    assert err <= 1e-12, err


@pytest.mark.parametrize("nu", [0.0, 0.05])
def test_locally_isothermal_block_viscous_fp64_exact(nu):
    """locally_isothermal::updated_conserved with face sound-speed tables: no
    library math functions are involved, so FP64 equals the oracle exactly."""
    import mara3p_amd as m3p
    n = 32
    rng = np.random.default_rng(21)
    uc, pe = _random_block(n, rng)
    cx, cy = c(rng.uniform(0.5, 2.0, (n + 1, n))), c(rng.uniform(0.5, 2.0, (n, n + 1)))
    dt, dl = 1e-3, 1.0 / n
    ref = np.zeros_like(uc)
    orc().orc_locally_isothermal_updated_conserved(P(uc), P(pe), n, n, dt, dl, P(cx), P(cy), nu, P(ref))
    out = torch.empty((3, n, n), dtype=torch.float64, device="cuda")
    d_uc, d_pe, d_cx, d_cy = dev(np.moveaxis(uc, -1, 0)), dev(np.moveaxis(pe, -1, 0)), dev(cx), dev(cy)
    m3p.check(m3p.lib().m3p_locally_isothermal_updated_conserved(ptr(d_uc), ptr(d_pe), ptr(out), n, dt, dl,
                                                                 ptr(d_cx), ptr(d_cy), nu, m3p.M3P_FP64, stream()))
    got = np.moveaxis(out.cpu().numpy(), 0, -1)
    assert np.array_equal(got, ref)


def test_minidisk_mesh_viscous_rk2_fp64():
    """Whole-mesh viscous minidisk stages (guard zones, corners included, read
    through the neighbour table as the reference's extend assembles them)
    against the oracle's mesh RK2."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=32, depth=2, domain_radius=2.0, softening_length=0.05, sink_radius=0.05,
                                 kinematic_viscosity=0.01)
    mesh, u0 = problems.minidisk(data, seed=4, noise=0.05)
    s = MinidiskSolver(data)
    s.set_state(u0)
    ref = to_reference_layout(u0).copy()
    t = 0.0
    for _ in range(3):
        dt = problems.minidisk_dt(data, t)
        s.step(dt)
        assert orc().orc_minidisk_mesh_rk2(P(ref), OrcMinidisk.from_data(data), t, dt, 1, 8) == 0
        t = t * 0.5 + (t + dt) * 0.5
    s.check_errors()
    err = rel_err(s.state(), to_device_layout(ref))
    assert err <= 1e-12, err


def test_minidisk_mesh_viscous_mixed_conserves_and_tracks_fp64():
    """Mixed viscous stages: with sinks and buffer off, mass is conserved to
    round-off; density stays within 1e-5 (L1, relative) of FP64 after 20 steps."""
    import math
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=64, depth=2, domain_radius=2.0, softening_length=0.05, sink_radius=0.05,
                                 kinematic_viscosity=0.02, sink_rate=0.0, buffer_rate=0.0)
    mesh, u0 = problems.minidisk(data, seed=5, noise=0.05)
    a, b = MinidiskSolver(data), MinidiskSolver(data, precision="mixed")
    a.set_state(u0)
    b.set_state(u0)
    t = 0.0
    for _ in range(20):
        dt = problems.minidisk_dt(data, t)
        a.step(dt)
        b.step(dt)
        t = t + dt
    a.check_errors()
    b.check_errors()
    m0 = math.fsum(u0[:, 0].ravel())
    ra, rb = a.state()[:, 0], b.state()[:, 0]
    assert abs(math.fsum(rb.ravel()) - m0) <= 1e-12 * m0
    assert abs(math.fsum(ra.ravel()) - m0) <= 1e-12 * m0
    assert np.mean(np.abs(ra - rb)) / np.mean(np.abs(ra)) < 1e-5
