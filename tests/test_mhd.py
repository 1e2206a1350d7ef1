"""Ideal MHD with constrained transport (mara::advance_mhd_ct,
src/scheme_mhd_v1.cpp:265-299): the oracle against the reference's golden
vectors and known-answer tests, then the HIP kernels against the oracle."""
import ctypes
import math
import os

import numpy as np
import pytest

import oracle
from oracle import P, orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
G = 5.0 / 3.0


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def golden():
    return np.load(os.path.join(GOLDEN, "mhd_ct.npz"))


# --------------------------------------------------------------------------- oracle pins
def test_mhd_riemann_golden():
    g = golden()
    for i in range(len(g["pl"])):
        f, e = np.zeros(5), np.zeros(3)
        orc().orc_mhd_riemann_hlle(P(c(g["pl"][i])), P(c(g["pr"][i])), float(g["bl"][i]), int(g["axis"][i]), G, P(f), P(e))
        assert np.array_equal(f, g["flux"][i]) and np.array_equal(e, g["efield"][i])


def test_mhd_advance_ct_golden():
    """One advance_mhd_ct call on an 8^3 periodic cube equals the reference exactly."""
    g = golden()
    N = int(g["N"])
    uo, o1, o2, o3, dt = (np.zeros_like(g["uc"]), np.zeros_like(g["bf1"]), np.zeros_like(g["bf2"]),
                          np.zeros_like(g["bf3"]), np.zeros(1))
    st = orc().orc_mhd_advance_ct(N, P(c(g["uc"])), P(c(g["bf1"])), P(c(g["bf2"])), P(c(g["bf3"])), 1.0 / N, 0.33, 1.5, G,
                                  P(uo), P(o1), P(o2), P(o3), P(dt))
    assert st == 0
    assert dt[0] == g["dt"][0]
    for a, b in ((uo, "uc_out"), (o1, "bf1_out"), (o2, "bf2_out"), (o3, "bf3_out")):
        assert np.array_equal(a, g[b])


@pytest.mark.parametrize("p0", [(1.0, 0, 0, 0, 1.0, 0, 0, 0), (1.5, 0, 0, 0, 1e3, 0, 0, 0),
                                (5.7, 1.0, 2.0, 3.0, 5.0, 0, 0, 0), (5.7, 1.0, 2.0, 3.0, 5.0, 3.0, 2.0, 1.0)])
def test_mhd_cons_prim_round_trip(p0):
    """test_mhd, src/physics_mhd.hpp:405-430 (1e-10 at gamma 5/3)."""
    p0 = c(p0)
    u, p1 = np.zeros(5), np.zeros(8)
    orc().orc_mhd_conserved_density(P(p0), G, P(u))
    assert orc().orc_mhd_recover_primitive(P(u), P(c(p0[5:])), G, P(p1)) == 0
    assert np.all(np.abs(p1 - p0) < 1e-10)


def test_mhd_recover_errors():
    assert orc().orc_mhd_recover_primitive(P(c([-1, 0, 0, 0, 1])), P(c([0, 0, 0])), G, P(np.zeros(8))) == 1
    assert orc().orc_mhd_recover_primitive(P(c([1, 0, 0, 0, 0.1])), P(c([1, 0, 0])), G, P(np.zeros(8))) == 2


def test_mhd_riemann_consistency():
    """Equal states: the HLLE flux is the physical flux and the face EMF is
    E = -v x B with the longitudinal field replaced."""
    p = c([1.2, 0.3, -0.2, 0.1, 0.9, 0.4, -0.5, 0.2])
    for ax in (1, 2, 3):
        f, e, fx = np.zeros(5), np.zeros(3), np.zeros(5)
        orc().orc_mhd_riemann_hlle(P(p), P(p), float(p[4 + ax]), ax, G, P(f), P(e))
        orc().orc_mhd_flux(P(p), ax, G, P(fx))
        assert np.allclose(f, fx, rtol=1e-13, atol=1e-14)
        E = -np.cross(p[1:4], p[5:8])
        n = np.eye(3)[ax - 1]
        assert np.allclose(e, E - np.dot(E, n) * n, atol=1e-14)


def test_oracle_ct_preserves_div_b():
    from mara3p_amd.problems import orszag_tang_3d
    N = 8
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=1, perturbation=0.05)
    uc = c(np.moveaxis(uc, 0, -1))
    b1, b2, b3 = c(b1), c(b2), c(b3)
    dts = np.zeros(2)
    for _ in range(3):
        assert orc().orc_mhd_rk2(N, P(uc), P(b1), P(b2), P(b3), 1.0 / N, 0.33, 1.5, G, P(dts)) == 0
    div = (b1[1:] - b1[:-1]) + (b2[:, 1:] - b2[:, :-1]) + (b3[:, :, 1:] - b3[:, :, :-1])
    assert np.abs(div).max() < 1e-12


# --------------------------------------------------------------------------- GPU
torch = pytest.importorskip("torch")
gpu = pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")


def _oracle_rk2(N, uc, b1, b2, b3, steps):
    u = c(np.moveaxis(uc, 0, -1))
    b1, b2, b3 = c(b1), c(b2), c(b3)
    dts = np.zeros(2)
    out = []
    for _ in range(steps):
        assert orc().orc_mhd_rk2(N, P(u), P(b1), P(b2), P(b3), 1.0 / N, 0.33, 1.5, G, P(dts)) == 0
        out.append(dts.copy())
    return np.moveaxis(u, -1, 0), b1, b2, b3, out


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.mark.gpu
@gpu
@pytest.mark.parametrize("N", [8, 12, 16, 72, 128])
def test_gpu_mhd_advance_ct_fp64_matches_oracle(N):
    """Two FP64 RK2 steps against the oracle: N = 12 runs the per-cell passes,
    8 and 16 the 8 x 8 column tiles, 128 the 8 x 32 tiles in 16-plane
    segments, 72 (N % 16 == 8) the 8 x 8 tiles in 8-plane segments; the
    tiled sizes update the face fields by bricks (2 x 4 x 8 cells, and
    4 x 4 x 32 at N = 128)."""
    from mara3p_amd.problems import orszag_tang_3d
    from mara3p_amd.solvers import MhdSolver
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=2, perturbation=0.05)
    s = MhdSolver(N)
    s.set_state(uc, b1, b2, b3)
    s.step()
    s.step()
    s.check_errors()
    got = s.state()
    ref = _oracle_rk2(N, uc, b1, b2, b3, 2)
    for a, b in zip(got, ref[:4]):
        assert _rel(a, b) <= 1e-12
    assert np.allclose(s.dts.cpu().numpy(), ref[4][-1], rtol=1e-13)


@pytest.mark.gpu
@gpu
def test_gpu_mhd_div_b_and_mixed():
    """Orszag-Tang at 32^3 for 10 steps: CT keeps max |div B| at round-off in both
    kernels; the mixed kernel's energy stays within 1e-4 (L1, relative) of FP64."""
    from mara3p_amd.problems import orszag_tang_3d
    from mara3p_amd.solvers import MhdSolver
    N = 32
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=3)
    a, b = MhdSolver(N), MhdSolver(N, precision="mixed")
    a.set_state(uc, b1, b2, b3)
    b.set_state(uc, b1, b2, b3)
    for _ in range(10):
        a.step()
        b.step()
    a.check_errors()
    scale = np.abs(b1).max() / (1.0 / N)
    assert a.max_div_b() < 1e-12 * scale
    assert b.max_div_b() < 1e-12 * scale
    ea, eb = a.state()[0][4], b.state()[0][4]
    assert np.mean(np.abs(ea - eb)) / np.mean(np.abs(ea)) < 1e-4


@pytest.mark.gpu
@gpu
@pytest.mark.parametrize("N", [12, 24, 32, 72])
def test_gpu_mhd_mixed_tracks_fp64(N):
    """Every mixed-precision MHD code path against the FP64 kernel, all four
    output arrays: N % 8 != 0 runs the per-cell passes (N = 12), the 8 x 8
    column tiles (N = 24), the 8 x 32 column tiles in one segment (N = 32) and
    in 8-plane segments (N = 72, N % 16 == 8).  Three RK2 steps from a
    perturbed Orszag-Tang state; each kernel takes its own CFL time step, so
    the bar (1e-5 of each array's max) covers FP32 fluxes and FP32 speeds."""
    from mara3p_amd.problems import orszag_tang_3d
    from mara3p_amd.solvers import MhdSolver
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=5, perturbation=0.05)
    a, b = MhdSolver(N), MhdSolver(N, precision="mixed")
    a.set_state(uc, b1, b2, b3)
    b.set_state(uc, b1, b2, b3)
    for _ in range(3):
        a.step()
        b.step()
    a.check_errors()
    b.check_errors()
    ga, gb = a.state(), b.state()
    for q in range(5):
        err = _rel(gb[0][q], ga[0][q])
        assert err <= 1e-5, (N, "U", q, err)
    for ax in range(1, 4):
        err = _rel(gb[ax], ga[ax])
        assert err <= 1e-5, (N, "B", ax, err)
    scale = np.abs(b1).max() / (1.0 / N)
    assert b.max_div_b() < 1e-12 * scale


@pytest.mark.gpu
@gpu
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_mhd_256_cubed_properties(precision):
    """The 3D MHD configuration at its full size, 256^3, by properties the
    oracle (serial) cannot reach there: from the z-independent Orszag-Tang
    state, two RK2 steps keep (1) max |div B| at round-off, (2) every z plane
    identical, (3) the point symmetry (x, y) -> (1 - x, 1 - y) of the initial
    state (density and energy even, momenta and fields odd), to 1e-13 of each
    array's max in FP64 and 1e-6 in the mixed kernel, and (4) the totals of
    mass, momentum and energy to round-off (periodic cube)."""
    from mara3p_amd.problems import orszag_tang_3d
    from mara3p_amd.solvers import MhdSolver
    N = 256
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=1, perturbation=0.0)
    s = MhdSolver(N, precision=precision)
    s.set_state(uc, b1, b2, b3)
    tot0 = [math.fsum(uc[q].ravel()) for q in range(5)]
    for _ in range(2):
        s.step()
    s.check_errors()
    scale = np.abs(b1).max() / (1.0 / N)
    assert s.max_div_b() < 1e-12 * scale
    u, f1, f2, f3 = s.state()
    assert np.array_equal(u, np.broadcast_to(u[:, :, :, :1], u.shape))
    assert np.abs(f3).max() == 0.0
    tol = 1e-13 if precision == "fp64" else 1e-6
    # reversing an axis maps cells i -> N - 1 - i and faces i -> N - i
    for q, sign in ((0, 1), (1, -1), (2, -1), (4, 1)):
        a = u[q, :, :, 0]
        assert _rel(sign * a[::-1, ::-1], a) <= tol, ("U", q)
    assert _rel(-f1[::-1, ::-1, 0], f1[:, :, 0]) <= tol
    assert _rel(-f2[::-1, ::-1, 0], f2[:, :, 0]) <= tol
    for q in (0, 4):
        t = math.fsum(u[q].ravel())
        assert abs(t - tot0[q]) <= 1e-13 * abs(tot0[q]), ("total", q)
    for q in (1, 2):
        t = math.fsum(u[q].ravel())
        assert abs(t - tot0[q]) <= 1e-13 * math.fsum(np.abs(uc[q]).ravel()), ("total", q)
