"""An ensemble of independent 1D Euler problems of uneven lengths advanced by
one persistent work-list launch per RK stage.  Each problem is the euler1d
chain of src/module_euler.cpp:107-161 with outflow ends (src/mara.cpp:123-143)
and cell spacing 1/L; the oracle runs each problem on its own."""
import numpy as np
import pytest

from oracle import P, orc
from mara3p_amd.mesh import EnsembleMesh1D
from mara3p_amd.problems import riemann_ensemble


def test_ensemble_mesh_layout():
    m = EnsembleMesh1D([128, 64, 256], 64)
    assert m.nblk == 7
    nbr = m.neighbours()
    assert nbr.tolist() == [[-1, 1], [0, -1], [-1, 2], [-1, 4], [3, 5], [4, 6], [5, -1]]
    assert np.allclose(m.cell_spacing(), [1 / 128] * 2 + [1 / 64] + [1 / 256] * 4)
    with pytest.raises(ValueError):
        EnsembleMesh1D([100], 64)


def test_riemann_ensemble_ranges():
    m, u = riemann_ensemble(count=512, block_size=64, seed=7)
    assert m.lengths.min() >= 512 and m.lengths.max() <= 8192
    assert np.all(m.lengths % 64 == 0)
    assert abs(m.lengths.mean() - 4352) < 400
    rho = u[:, 0]
    assert rho.min() >= 0.1 and rho.max() <= 10.0
    assert np.all(u[:, 1:4] == 0.0)
    # the same seed gives the same ensemble
    m2, u2 = riemann_ensemble(count=512, block_size=64, seed=7)
    assert np.array_equal(u, u2)


def _oracle_problem(u_dev_blocks, L, n, dt, steps, gamma=1.4):
    """Advance one problem with the oracle's euler1d chain (outflow ends)."""
    u = np.ascontiguousarray(np.moveaxis(u_dev_blocks, 1, -1))     # [nb][n][5]
    assert orc().orc_euler1d_mesh_rk2(P(u), L // n, n, dt, 1.0 / L, 1.5, gamma, steps, 1) == 0
    return np.moveaxis(u, -1, 1)


def test_oracle_ensemble_member_conserves_mass_before_waves_reach_ends():
    m, u = riemann_ensemble(count=3, block_size=64, lmin=512, lmax=1024, seed=2)
    for k in range(3):
        L = int(m.lengths[k])
        uk = u[m.problem_slice(k)]
        out = _oracle_problem(uk, L, 64, 0.1 / L, 20)
        assert abs(out[:, 0].sum() - uk[:, 0].sum()) <= 1e-12 * uk[:, 0].sum()


torch = pytest.importorskip("torch")
gpu = pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")


@pytest.mark.gpu
@gpu
@pytest.mark.parametrize("count,lmin,lmax,seed,steps", [(24, 128, 2048, 3, 4), (512, 512, 8192, 1, 1)])
def test_gpu_ensemble_fp64_matches_oracle_per_problem(count, lmin, lmax, seed, steps):
    """Every problem of the ensemble equals the oracle's own chain run bit for
    bit; (512, 512, 8192, 1) is the ensemble configuration at its full size."""
    from mara3p_amd.solvers import Euler1DEnsembleSolver
    n = 64
    m, u = riemann_ensemble(count=count, block_size=n, lmin=lmin, lmax=lmax, seed=seed)
    s = Euler1DEnsembleSolver(m)
    s.set_state(u)
    dt = 0.05 / 8192
    for _ in range(steps):
        s.step(dt)
    s.check_errors()
    got = s.state()
    for k in range(count):
        L = int(m.lengths[k])
        ref = _oracle_problem(u[m.problem_slice(k)], L, n, dt, steps)
        assert np.array_equal(got[m.problem_slice(k)], ref), k


@pytest.mark.gpu
@gpu
def test_gpu_ensemble_cfl_and_mixed():
    """The full 512-problem ensemble: the device CFL dt is cfl * min over blocks
    of dl_b / max speed; the mixed kernel stays within 1e-5 (L1 of rho,
    relative) of FP64 after 50 CFL-limited steps and both conserve mass until
    waves reach the outflow ends."""
    from mara3p_amd.solvers import Euler1DEnsembleSolver
    m, u = riemann_ensemble(count=512, block_size=64, seed=1)
    a, b = Euler1DEnsembleSolver(m), Euler1DEnsembleSolver(m, precision="mixed")
    a.set_state(u)
    b.set_state(u)
    dt = float(a.cfl_dt(0.4).item())
    D = u[:, 0]
    v = u[:, 1] / D
    pg = (u[:, 4] - 0.5 * (v * u[:, 1])) * (1.4 - 1.0)
    speed = np.abs(v) + np.sqrt(1.4 * pg / D)
    want = 0.4 * np.min(m.cell_spacing()[:, None] / speed)
    assert abs(dt - want) <= 1e-14 * want
    for _ in range(50):
        a.step_cfl(0.4)
        b.step_cfl(0.4)
    a.check_errors()
    b.check_errors()
    ra, rb = a.state()[:, 0], b.state()[:, 0]
    assert np.mean(np.abs(ra - rb)) / np.mean(np.abs(ra)) < 1e-5
    dl = m.cell_spacing()[:, None]
    mass0 = np.sum(u[:, 0] * dl)
    assert abs(np.sum(ra * dl) - mass0) <= 1e-12 * mass0
    assert abs(np.sum(rb * dl) - mass0) <= 1e-12 * mass0
