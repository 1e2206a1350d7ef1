"""GPU parity: the HIP kernels, called through the C ABI, against the oracle.

FP64 Euler (HLLE and HLLC) must equal the oracle exactly (== on every value).
FP64 minidisk must agree to 1e-12 relative per zone per step (normalised by
the component's max |u|): its sources call exp/pow/tanh, whose device (ocml)
and host (glibc) implementations may differ by an ulp.  The mixed kernel is
checked against FP64 at a stated tolerance and for exact conservation.
"""
import ctypes
import math

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")]

import oracle  # noqa: E402
from oracle import OrcMinidisk, P, orc  # noqa: E402


def _m3p():
    import mara3p_amd
    return mara3p_amd


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


def smooth_euler_block(n, rng, ext=True):
    """Random smooth positive primitive field on an (n+4)^2 extended block, conserved interior."""
    m = n + 4
    x = np.arange(m)[:, None] / m
    y = np.arange(m)[None, :] / m
    k = rng.uniform(1, 3, 6)
    rho = 1.0 + 0.3 * np.sin(2 * np.pi * (k[0] * x + k[1] * y)) + 0.05 * rng.random((m, m))
    v1 = 0.4 * np.cos(2 * np.pi * k[2] * y) + 0.05 * rng.standard_normal((m, m))
    v2 = 0.3 * np.sin(2 * np.pi * k[3] * x) + 0.05 * rng.standard_normal((m, m))
    v3 = 0.1 * rng.standard_normal((m, m))
    p = 1.0 + 0.2 * np.cos(2 * np.pi * (k[4] * x - k[5] * y)) + 0.05 * rng.random((m, m))
    pe = np.stack([rho, v1, v2, v3, p], -1)
    uc = np.zeros((n, n, 5))
    for i in range(n):
        for j in range(n):
            orc().orc_euler_conserved_density(P(c(pe[i + 2, j + 2])), 1.4, P(uc[i, j]))
    uc *= 1.0 + 1e-3 * rng.standard_normal((n, n, 1))
    return c(uc), c(pe)


# --------------------------------------------------------------------------- per-block API
@pytest.mark.parametrize("n", [16, 64, 256])
@pytest.mark.parametrize("hllc", [0, 1])
def test_euler2d_updated_conserved_fp64_exact(n, hllc):
    m3p = _m3p()
    rng = np.random.default_rng(n + 10 * hllc)
    uc, pe = smooth_euler_block(n, rng)
    dt, dl = 2e-3, 1.0 / n
    ref = np.zeros_like(uc)
    orc().orc_euler2d_updated_conserved(P(uc), P(pe), n, n, dt, dl, 1.5, 1.4, hllc, P(ref))
    out = torch.empty((5, n, n), dtype=torch.float64, device="cuda")
    d_uc, d_pe = dev(np.moveaxis(uc, -1, 0)), dev(np.moveaxis(pe, -1, 0))   # held until the kernel has run
    m3p.check(m3p.lib().m3p_euler2d_updated_conserved(ptr(d_uc), ptr(d_pe), ptr(out), n, dt, dl, 1.5, 1.4, hllc,
                                                      m3p.M3P_FP64, stream()))
    got = np.moveaxis(out.cpu().numpy(), 0, -1)
    assert np.array_equal(got, ref)


def test_euler1d_updated_conserved_fp64_exact():
    m3p = _m3p()
    rng = np.random.default_rng(3)
    n = 64
    for lo, hi in ((0, 0), (1, 0), (0, 1), (1, 1)):
        pe = np.stack([1 + 0.3 * rng.random(n + 4), 0.3 * rng.standard_normal(n + 4), 0.1 * rng.standard_normal(n + 4),
                       0 * rng.random(n + 4), 1 + 0.3 * rng.random(n + 4)], -1)
        uc = np.zeros((n, 5))
        for i in range(n):
            orc().orc_euler_conserved_density(P(c(pe[i + 2])), 1.4, P(uc[i]))
        uc, pe = c(uc), c(pe)
        ref = np.zeros_like(uc)
        orc().orc_euler1d_updated_conserved(P(uc), P(pe), n, 1e-3, 1.0 / n, 1.5, 1.4, lo, hi, P(ref))
        out = torch.empty((5, n), dtype=torch.float64, device="cuda")
        d_uc, d_pe = dev(uc.T), dev(pe.T)
        m3p.check(m3p.lib().m3p_euler1d_updated_conserved(ptr(d_uc), ptr(d_pe), ptr(out), n, 1e-3, 1.0 / n,
                                                          1.5, 1.4, lo, hi, m3p.M3P_FP64, stream()))
        assert np.array_equal(out.cpu().numpy().T, ref)


@pytest.mark.parametrize("n,level,bi,bj", [(16, 2, 1, 2), (64, 1, 0, 1), (256, 4, 7, 9)])
def test_minidisk_updated_conserved_fp64(n, level, bi, bj):
    m3p = _m3p()
    rng = np.random.default_rng(level)
    data = m3p.MinidiskData.defaults(block_size=n, depth=level, domain_radius=8.0, softening_length=0.05,
                                     sink_radius=0.05, eccentricity=0.1, mass_ratio=0.8)
    sd = OrcMinidisk.from_data(data)
    ui = np.zeros((n, n, 3))
    orc().orc_minidisk_initial_conserved(ctypes.byref(sd), level, bi, bj, P(ui))
    uc = c(ui * np.exp(0.05 * rng.standard_normal((n, n, 1))))
    pe = np.zeros((n + 4, n + 4, 3))
    pe[..., 0] = rng.uniform(0.5, 1.5, (n + 4, n + 4))
    pe[..., 1:] = rng.standard_normal((n + 4, n + 4, 2))
    pe = c(pe)
    ref = np.zeros_like(uc)
    orc().orc_minidisk_updated_conserved(P(uc), P(pe), ctypes.byref(sd), level, bi, bj, 0.9, 1e-3, P(ref))
    out = torch.empty((3, n, n), dtype=torch.float64, device="cuda")
    d_uc, d_pe = dev(np.moveaxis(uc, -1, 0)), dev(np.moveaxis(pe, -1, 0))
    m3p.check(m3p.lib().m3p_minidisk_updated_conserved(ptr(d_uc), ptr(d_pe), ptr(out), level, bi, bj, 0.9, 1e-3,
                                                       ctypes.byref(data), m3p.M3P_FP64, stream()))
    got = np.moveaxis(out.cpu().numpy(), 0, -1)
    err = rel_err(got[None].transpose(0, 3, 1, 2), ref[None].transpose(0, 3, 1, 2))
    assert err <= 1e-12, err


# --------------------------------------------------------------------------- whole-mesh stages
@pytest.mark.parametrize("depth,n,hllc", [(1, 64, 0), (2, 32, 0), (1, 64, 1), (2, 256, 0), (4, 256, 1)])
def test_euler2d_mesh_rk2_fp64_exact(depth, n, hllc):
    """(4, 256, 1) is the Kelvin-Helmholtz configuration at its full size,
    4096^2 with HLLC, one RK2 step."""
    from mara3p_amd import problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import Euler2DSolver
    mesh, u0 = problems.kelvin_helmholtz(depth=depth, block_size=n, seed=depth + n)
    s = Euler2DSolver(depth, n, gamma=5.0 / 3.0, plm_theta=1.5, hllc=bool(hllc))
    s.set_state(u0)
    dt = 0.4 * s.dl / 2.0
    steps = 3 if n < 256 else 1
    for _ in range(steps):
        s.step(dt)
    s.check_errors()
    ref = to_reference_layout(u0).copy()
    assert orc().orc_euler2d_mesh_rk2(P(ref), mesh.nb, n, dt, s.dl, 1.5, 5.0 / 3.0, hllc, steps, 8) == 0
    assert np.array_equal(s.state(), to_device_layout(ref))


def _with_tile(kind, precision, tj, h):
    """Set a stage tile shape; returns a restore callback."""
    from mara3p_amd._lib import lib
    otj, oh = ctypes.c_int(0), ctypes.c_int(0)
    lib().m3p_get_stage2d_tile(kind, precision, ctypes.byref(otj), ctypes.byref(oh))
    assert lib().m3p_set_stage2d_tile(kind, precision, tj, h) == 0
    return lambda: lib().m3p_set_stage2d_tile(kind, precision, otj.value, oh.value)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_tile_shape_gives_the_same_stage(precision):
    """Every compiled tile shape computes the same face fluxes in the same
    order, so a stage is bitwise independent of the tiling; FP64 also equals
    the oracle."""
    from mara3p_amd import problems
    from mara3p_amd._lib import M3P_KIND_EULER2D, PRECISIONS
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import Euler2DSolver, TILE_VARIANTS
    depth, n = 1, 256
    mesh, u0 = problems.kelvin_helmholtz(depth=depth, block_size=n, seed=11)
    s = Euler2DSolver(depth, n, precision=precision)
    dt = 0.4 * s.dl / 2.0
    results = {}
    for tj, h in TILE_VARIANTS:
        restore = _with_tile(M3P_KIND_EULER2D, PRECISIONS[precision], tj, h)
        try:
            s.set_state(u0)
            s.step(dt)
            s.check_errors()
            results[(tj, h)] = s.state()
        finally:
            restore()
    first = next(iter(results.values()))
    for k, v in results.items():
        assert np.array_equal(v, first), k
    if precision == "fp64":
        ref = to_reference_layout(u0).copy()
        assert orc().orc_euler2d_mesh_rk2(P(ref), mesh.nb, n, dt, s.dl, 1.5, 5.0 / 3.0, 0, 1, 8) == 0
        assert np.array_equal(first, to_device_layout(ref))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_every_tile_shape_gives_the_same_minidisk_step(precision):
    """The minidisk scheme's stage kernels (its first stage has instantiations
    of its own, fetching two rows ahead) give bitwise the same RK2 step for
    every compiled tile shape; FP64 is within 1e-12 of the oracle."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd._lib import M3P_KIND_MINIDISK, PRECISIONS
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import MinidiskSolver, TILE_VARIANTS
    data = MinidiskData.defaults(block_size=256, depth=1, domain_radius=2.0, softening_length=0.05,
                                 sink_radius=0.05)
    mesh, u0 = problems.minidisk(data, seed=13)
    s = MinidiskSolver(data, precision)
    dt = problems.minidisk_dt(data, 0.0)
    results = {}
    for tj, h in TILE_VARIANTS:
        restore = _with_tile(M3P_KIND_MINIDISK, PRECISIONS[precision], tj, h)
        try:
            s.set_state(u0)
            s.time = 0.0          # the binary's position follows the solution time
            s.step(dt)
            s.check_errors()
            results[(tj, h)] = s.state()
        finally:
            restore()
    first = next(iter(results.values()))
    for k, v in results.items():
        assert np.array_equal(v, first), k
    if precision == "fp64":
        ref = to_reference_layout(u0).copy()
        orc().orc_minidisk_mesh_rk2(P(ref), OrcMinidisk.from_data(data), 0.0, dt, 1, 8)
        err = rel_err(first, to_device_layout(ref))
        assert err <= 1e-12, err


def test_minidisk_mesh_rk2_fp64():
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=64, depth=2, domain_radius=2.0, softening_length=0.05, sink_radius=0.05)
    mesh, u0 = problems.minidisk(data, seed=9)
    s = MinidiskSolver(data)
    s.set_state(u0)
    ref = to_reference_layout(u0).copy()
    t = 0.0
    for _ in range(3):
        dt = problems.minidisk_dt(data, t)
        s.step(dt)
        orc().orc_minidisk_mesh_rk2(P(ref), OrcMinidisk.from_data(data), t, dt, 1, 8)
        t = t * 0.5 + (t + dt) * 0.5
        err = rel_err(s.state(), to_device_layout(ref))
        assert err <= 1e-12, err


def test_euler1d_sod_mesh_fp64_exact():
    from mara3p_amd import problems
    from mara3p_amd.mesh import to_reference_layout
    from mara3p_amd.solvers import Euler1DSolver
    nblk, n = 16, 64
    u0 = problems.sod(nblk, n, 1.4)
    s = Euler1DSolver(nblk, n, gamma=1.4, plm_theta=1.5)
    s.set_state(u0)
    dt = 0.4 * s.dl / 2.2
    for _ in range(40):
        s.step(dt)
    s.check_errors()
    ref = c(u0.transpose(0, 2, 1))
    orc().orc_euler1d_mesh_rk2(P(ref), nblk, n, dt, s.dl, 1.5, 1.4, 40, 1)
    assert np.array_equal(s.state(), ref.transpose(0, 2, 1))


@pytest.mark.parametrize("n,nblk", [(16, 24), (32, 12), (128, 6), (256, 4), (512, 3)])
def test_euler1d_mesh_every_block_size(n, nblk):
    """Every compiled 1D tile width (16 .. 256 zones; 512-zone blocks have two
    tiles each, 256 zones span four waves): a perturbed Sod chain, five RK2
    steps; FP64 equals the oracle bit for bit, mixed is within 1e-5 of it."""
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler1DSolver
    rng = np.random.default_rng(n)
    u0 = problems.sod(nblk, n, 1.4).copy()
    u0[:, 0] *= 1.0 + 0.01 * rng.random((nblk, n))
    u0[:, 1] = 0.05 * rng.standard_normal((nblk, n)) * u0[:, 0]
    u0[:, 4] += 0.5 * u0[:, 1] ** 2 / u0[:, 0]
    steps = 5
    out = {}
    for precision in ("fp64", "mixed"):
        s = Euler1DSolver(nblk, n, gamma=1.4, plm_theta=1.5, precision=precision)
        s.set_state(u0)
        dt = 0.4 * s.dl / 2.2
        for _ in range(steps):
            s.step(dt)
        s.check_errors()
        out[precision] = s.state()
    ref = c(u0.transpose(0, 2, 1))
    orc().orc_euler1d_mesh_rk2(P(ref), nblk, n, dt, s.dl, 1.5, 1.4, steps, 1)
    assert np.array_equal(out["fp64"], ref.transpose(0, 2, 1))
    err = rel_err(out["mixed"], out["fp64"])
    assert err <= 1e-5, err


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sod_to_t_0_2(precision):
    """The Sod configuration (1024 zones, gamma 1.4, PLM 1.5, HLLE, RK2, CFL
    0.4 by the signal-speed bound 2.2, outflow ends) integrated to t = 0.2:
    FP64 equals the oracle over the same time steps bit for bit; both
    precisions are within L1 3e-3 of the exact density (the oracle's own
    bar) and the mixed density within 1e-6 (L1) of FP64."""
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler1DSolver
    nblk, n, g = 16, 64, 1.4
    u0 = problems.sod(nblk, n, g)
    dl = 1.0 / (nblk * n)
    nsteps = int(np.ceil(0.2 / (0.4 * dl / 2.2)))
    dt = 0.2 / nsteps
    s = Euler1DSolver(nblk, n, gamma=g, plm_theta=1.5, precision=precision)
    s.set_state(u0)
    for _ in range(nsteps):
        s.step(dt)
    s.check_errors()
    ref = c(u0.transpose(0, 2, 1))
    assert orc().orc_euler1d_mesh_rk2(P(ref), nblk, n, dt, dl, 1.5, g, nsteps, 4) == 0
    ref = ref.transpose(0, 2, 1)
    got = s.state()
    if precision == "fp64":
        assert np.array_equal(got, ref)
    else:
        assert np.mean(np.abs(got[:, 0] - ref[:, 0])) < 1e-6
    rho = got[:, 0].reshape(-1)
    xc = (np.arange(nblk * n) + 0.5) * dl - 0.5
    exact = np.zeros(3)
    left, right = c([1.0, 0.0, 1.0]), c([0.125, 0.0, 0.1])
    ex = np.zeros_like(xc)
    for i, xi in enumerate(xc):
        orc().orc_euler_exact_riemann(P(left), P(right), g, xi / 0.2, P(exact))
        ex[i] = exact[0]
    assert np.mean(np.abs(rho - ex)) < 3e-3


def test_minidisk_full_size_rk2_fp64():
    """The minidisk configuration at its full size (16 x 16 blocks of 256^2 on
    [-8, 8]^2), one RK2 step: FP64 within 1e-12 of the oracle."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05,
                                 sink_radius=0.05, mach_number=10.0)
    mesh, u0 = problems.minidisk(data, seed=1, noise=0.01)
    s = MinidiskSolver(data)
    s.set_state(u0)
    dt = problems.minidisk_dt(data, 0.0)
    s.step(dt)
    s.check_errors()
    ref = to_reference_layout(u0).copy()
    assert orc().orc_minidisk_mesh_rk2(P(ref), OrcMinidisk.from_data(data), 0.0, dt, 1, 16) == 0
    err = rel_err(s.state(), to_device_layout(ref))
    assert err <= 1e-12, err


def per_zone_rel_err(got, ref, floor_frac):
    """max over zones and components of |got - ref| / max(|ref|, floor), with
    floor = floor_frac * the component's max |ref| (component axis = 1)."""
    axes = tuple(i for i in range(ref.ndim) if i != 1)
    floor = floor_frac * np.abs(ref).max(axis=axes, keepdims=True)
    return float((np.abs(got - ref) / np.maximum(np.abs(ref), np.maximum(floor, 1e-300))).max())


def test_minidisk_full_size_per_zone_fp64():
    """The minidisk configuration at full size, one RK2 step, FP64 against the
    oracle zone by zone: |got - ref| / max(|ref|, floor) with the floor at
    1e-6 of the component's max (the momenta change sign across the mesh,
    where a zone's value is a small remainder of O(1) terms); the bar is
    1e-12, the same as the global-max norm of the test above."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05,
                                 sink_radius=0.05, mach_number=10.0)
    mesh, u0 = problems.minidisk(data, seed=1, noise=0.01)
    s = MinidiskSolver(data)
    s.set_state(u0)
    dt = problems.minidisk_dt(data, 0.0)
    s.step(dt)
    s.check_errors()
    ref = to_reference_layout(u0).copy()
    assert orc().orc_minidisk_mesh_rk2(P(ref), OrcMinidisk.from_data(data), 0.0, dt, 1, 16) == 0
    ref = to_device_layout(ref)
    got = s.state()
    figures = {f: per_zone_rel_err(got, ref, f) for f in (1e-3, 1e-6, 1e-9)}
    print("minidisk FP64 per-zone relative error by floor:", figures)
    assert figures[1e-6] <= 1e-12, figures


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_minidisk_full_size_mass_conservation(precision):
    """The minidisk configuration at full size (16 x 16 blocks of 256^2) with
    the sink and buffer rates at zero, 1000 RK2 steps: total mass conserved
    to <= 1e-13 relative in both kernels (each face flux is computed once and
    differenced in FP64, the mesh is periodic)."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05,
                                 sink_radius=0.05, mach_number=10.0, sink_rate=0.0, buffer_rate=0.0)
    _, u0 = problems.minidisk(data, seed=1, noise=0.01)
    s = MinidiskSolver(data, precision)
    s.set_state(u0)
    m0 = math.fsum(u0[:, 0].ravel())
    for _ in range(1000):
        s.step(problems.minidisk_dt(data, s.time))
    s.check_errors()
    u = s.state()
    assert np.all(np.isfinite(u))
    m1 = math.fsum(u[:, 0].ravel())
    drift = abs(m1 - m0) / m0
    print("minidisk %s mass drift after 1000 steps: %.3e" % (precision, drift))
    assert drift <= 1e-13, drift


# --------------------------------------------------------------------------- mixed precision
def test_minidisk_mixed_tracks_fp64():
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=64, depth=2, domain_radius=2.0, softening_length=0.05, sink_radius=0.05)
    _, u0 = problems.minidisk(data, seed=2)
    a, b = MinidiskSolver(data, "fp64"), MinidiskSolver(data, "mixed")
    a.set_state(u0)
    b.set_state(u0)
    t = 0.0
    for _ in range(20):
        dt = problems.minidisk_dt(data, t)
        a.step(dt)
        b.step(dt)
        t = a.time
    ua, ub = a.state(), b.state()
    l1 = np.mean(np.abs(ua[:, 0] - ub[:, 0]))
    assert l1 < 1e-5, l1
    assert rel_err(ub, ua) < 1e-3


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_minidisk_mass_conservation(precision):
    """With sink and buffer rates zero the mass update is a pure flux difference:
    total mass is conserved to FP64 round-off (<= 1e-13 relative) in both kernels."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=64, depth=2, domain_radius=2.0, softening_length=0.05,
                                 sink_radius=0.05, sink_rate=0.0, buffer_rate=0.0)
    _, u0 = problems.minidisk(data, seed=4)
    s = MinidiskSolver(data, precision)
    s.set_state(u0)
    m0 = math.fsum(u0[:, 0].ravel())
    t = 0.0
    for _ in range(50):
        s.step(problems.minidisk_dt(data, t))
        t = s.time
    m1 = math.fsum(s.state()[:, 0].ravel())
    assert abs(m1 - m0) / m0 <= 1e-13
    assert abs(s.conserved_totals()[0] - m1) / m0 <= 1e-14


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_euler2d_conservation(precision):
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=2, block_size=64, seed=8)
    s = Euler2DSolver(2, 64, precision=precision)
    s.set_state(u0)
    tot0 = np.array([math.fsum(u0[:, q].ravel()) for q in range(5)])
    for _ in range(20):
        s.step(0.4 * s.dl / 2.0)
    u = s.state()
    tot = np.array([math.fsum(u[:, q].ravel()) for q in range(5)])
    scale = np.array([math.fsum(np.abs(u0[:, q]).ravel()) for q in range(5)])
    assert np.all(np.abs(tot - tot0) <= 1e-13 * scale)


# --------------------------------------------------------------------------- reductions / errors
def test_cfl_dt_matches_numpy():
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=1, block_size=64, seed=2)
    s = Euler2DSolver(1, 64)
    s.set_state(u0)
    dt = float(s.cfl_dt(0.4).item())
    D = u0[:, 0]
    v1, v2, v3 = u0[:, 1] / D, u0[:, 2] / D, u0[:, 3] / D
    pg = (u0[:, 4] - 0.5 * (v1 * u0[:, 1] + v2 * u0[:, 2] + v3 * u0[:, 3])) * (5.0 / 3.0 - 1)
    vmax = (np.maximum(np.abs(v1), np.abs(v2)) + np.sqrt(5.0 / 3.0 * pg / D)).max()
    assert abs(dt - 0.4 * s.dl / vmax) <= 1e-14 * dt


def test_totals_match_fsum():
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=2, block_size=64, seed=5)
    s = Euler2DSolver(2, 64)
    s.set_state(u0)
    tot = s.conserved_totals()
    for q in range(5):
        ref = math.fsum(u0[:, q].ravel())
        assert abs(tot[q] - ref) <= 1e-14 * max(1.0, math.fsum(np.abs(u0[:, q]).ravel()))


def test_negative_density_raises_reference_error():
    from mara3p_amd import NegativeDensityError, problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=1, block_size=64, seed=2)
    u0 = u0.copy()
    u0[2, 0, 10, 20] = -1.0
    s = Euler2DSolver(1, 64)
    s.set_state(u0)
    s.step(1e-4)
    with pytest.raises(NegativeDensityError):
        s.check_errors()
    s.check_errors()   # latch is cleared


def test_dt_from_device_matches_host_dt():
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=1, block_size=64, seed=6)
    a, b = Euler2DSolver(1, 64), Euler2DSolver(1, 64)
    a.set_state(u0)
    b.set_state(u0)
    dt = float(a.cfl_dt(0.4).item())
    a.step_cfl(0.4)
    b.step(dt)
    assert np.array_equal(a.state(), b.state())
