"""Device-resident solvers: the whole mesh lives in HBM for the run; each RK2
step is two fused stage launches (control::advance_runge_kutta with
weighted_sum_tree, src/app_control.hpp:169-195, src/mesh_amr.hpp:455-467).

PyTorch provides device memory and the stream; all arithmetic is in the HIP
library.
"""
import ctypes

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, MinidiskData, PRECISIONS
from .amr import QuadTree
from .mesh import ChainMesh1D, EnsembleMesh1D, UniformMesh2D


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


# The compiled stage-kernel tile shapes (columns x rows), as in capi.hip.
TILE_VARIANTS = ((256, 128), (256, 64), (256, 32), (256, 16), (128, 64), (128, 32), (64, 64), (64, 32), (64, 16),
                 (32, 32), (16, 16))


def autotune_tiles(kind, precision, n, run, reps=5):
    """Time run() (one stage launch on the caller's arrays) under every tile
    shape that fits block size n, keep the fastest for (kind, precision);
    returns (tj, h, ms)."""
    best = None
    for tj, h in TILE_VARIANTS:
        if n % tj or n % h:
            continue
        check(lib().m3p_set_stage2d_tile(kind, precision, tj, h))
        run()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            run()
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / reps
        if best is None or ms < best[2]:
            best = (tj, h, ms)
    check(lib().m3p_set_stage2d_tile(kind, precision, best[0], best[1]))
    return best


class _StateSolver:
    nc = None

    def _alloc(self, shape, device):
        self.device = torch.device(device)
        self.u = torch.zeros(shape, dtype=torch.float64, device=self.device)
        self.u1 = torch.empty_like(self.u)
        self.u2 = torch.empty_like(self.u)
        self.totals = torch.zeros(self.nc, dtype=torch.float64, device=self.device)

    def set_state(self, u):
        self.u.copy_(torch.as_tensor(np.ascontiguousarray(u), dtype=torch.float64))

    def state(self):
        return self.u.cpu().numpy()

    def conserved_totals(self):
        zpb = int(np.prod(self.u.shape[2:]))
        check(lib().m3p_conserved_totals(_ptr(self.u), self.u.shape[0], self.nc, zpb, _ptr(self.totals), _stream()))
        return self.totals.cpu().numpy()

    def check_errors(self):
        _lib.check_kernel_errors(torch.cuda.current_stream().cuda_stream)


class Euler2DSolver(_StateSolver):
    """euler2d on a uniform periodic quad-tree level (module_euler.cpp:263-323)."""
    nc = 5

    def __init__(self, depth, block_size, gamma=5.0 / 3.0, plm_theta=1.5, precision="fp64",
                 hllc=False, domain_size=1.0, device="cuda"):
        self.mesh = UniformMesh2D(depth, block_size)
        self.gamma, self.theta, self.hllc = gamma, plm_theta, int(hllc)
        self.precision = PRECISIONS[precision]
        # cell_spacing, src/mesh_geometry.cpp:168-171
        self.dl = domain_size / float(block_size) / float(1 << depth)
        self._alloc((self.mesh.nblk, 5, block_size, block_size), device)
        self.nbr = torch.as_tensor(self.mesh.neighbours(), device=self.device)
        self.dt_dev = torch.zeros(1, dtype=torch.float64, device=self.device)

    def stage(self, u, u0, out, dt, rk_b=0.5, dt_dev=None, cfl=None, dt_next=None):
        """One stage; with cfl and dt_next (device double) the CFL time step of
        `out` is reduced in the same launch and written to dt_next."""
        if cfl is not None:
            check(lib().m3p_euler2d_stage_cfl(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), self.mesh.nblk,
                                              self.mesh.n, dt, _ptr(dt_dev), self.dl, rk_b, self.theta, self.gamma,
                                              self.hllc, cfl, _ptr(dt_next), self.precision, _stream()))
            return
        check(lib().m3p_euler2d_stage(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), self.mesh.nblk, self.mesh.n,
                                      dt, _ptr(dt_dev), self.dl, rk_b, self.theta, self.gamma, self.hllc,
                                      self.precision, _stream()))

    def step(self, dt):
        self.stage(self.u, None, self.u1, dt)
        self.stage(self.u1, self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u

    def rk2_step(self, u, out, dt, rk_b=0.5, dt_dev=None, cfl=None, dt_next=None):
        """One whole RK2 step u -> out in one launch (the fused kernel, mixed
        precision: the first stage's state is never stored); equal to the two
        stages.  With cfl and dt_next the CFL time step of `out` is reduced in
        the same launch."""
        if cfl is not None:
            check(lib().m3p_euler2d_rk2_step_cfl(_ptr(u), _ptr(out), _ptr(self.nbr), self.mesh.nblk, self.mesh.n,
                                                 dt, _ptr(dt_dev), self.dl, rk_b, self.theta, self.gamma, self.hllc,
                                                 cfl, _ptr(dt_next), self.precision, _stream()))
            return
        check(lib().m3p_euler2d_rk2_step(_ptr(u), _ptr(out), _ptr(self.nbr), self.mesh.nblk, self.mesh.n, dt,
                                         _ptr(dt_dev), self.dl, rk_b, self.theta, self.gamma, self.hllc,
                                         self.precision, _stream()))

    def step_fused(self, dt):
        """step(dt) through the fused RK2 kernel (mixed precision)."""
        self.rk2_step(self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u

    def cfl_dt(self, cfl):
        """CFL dt from the current state into self.dt_dev (device) and returned."""
        zpb = self.mesh.n * self.mesh.n
        check(lib().m3p_euler_cfl_dt(_ptr(self.u), self.mesh.nblk, zpb, self.gamma, cfl, self.dl,
                                     _ptr(self.dt_dev), _stream()))
        return self.dt_dev

    def autotune(self, reps=5):
        """Pick the fastest compiled tile shape for this mesh (blocking); returns (tj, h, ms)."""
        tj, h, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0)
        check(lib().m3p_autotune_euler2d(_ptr(self.u), _ptr(self.u1), _ptr(self.nbr), self.mesh.nblk, self.mesh.n,
                                         self.dl, self.gamma, self.precision, reps, ctypes.byref(tj), ctypes.byref(h),
                                         ctypes.byref(ms), _stream()))
        return tj.value, h.value, ms.value

    def step_cfl(self, cfl):
        """One RK2 step with dt = cfl dl / max(|v| + c_s) kept on the device."""
        dt_dev = self.cfl_dt(cfl)
        self.stage(self.u, None, self.u1, 0.0, dt_dev=dt_dev)
        self.stage(self.u1, self.u, self.u2, 0.0, dt_dev=dt_dev)
        self.u, self.u2 = self.u2, self.u


class MinidiskSolver(_StateSolver):
    """minidisk::updated_conserved on every block, RK2 (problems/minidisk_scheme.cpp:294-355)."""
    nc = 3

    def __init__(self, data: MinidiskData, precision="fp64", device="cuda"):
        self.data = data
        self.mesh = UniformMesh2D(data.depth, data.block_size)
        self.precision = PRECISIONS[precision]
        n = data.block_size
        self._alloc((self.mesh.nblk, 3, n, n), device)
        self.nbr = torch.as_tensor(self.mesh.neighbours(), device=self.device)
        self.bcoord = torch.as_tensor(self.mesh.block_coordinates(), device=self.device)
        self.time = 0.0
        # device-clocked steps: {t, dt} and the binary's state at the two stage times
        self.clock = torch.zeros(2, dtype=torch.float64, device=self.device)
        self.orbit = torch.zeros(12, dtype=torch.float64, device=self.device)

    def stage(self, u, u0, out, time, dt, rk_b=0.5):
        check(lib().m3p_minidisk_stage(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), _ptr(self.bcoord),
                                       self.mesh.nblk, time, dt, rk_b, ctypes.byref(self.data),
                                       self.precision, _stream()))

    def stage_dev(self, u, u0, out, k, rk_b=0.5):
        """A stage whose dt and binary state come from device memory (the
        clock's dt, the orbit of stage k of the step)."""
        orbit = ctypes.c_void_p(self.orbit.data_ptr() + 6 * 8 * k)
        dt = ctypes.c_void_p(self.clock.data_ptr() + 8)
        check(lib().m3p_minidisk_stage_dev(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), _ptr(self.bcoord),
                                           self.mesh.nblk, orbit, dt, rk_b, ctypes.byref(self.data),
                                           self.precision, _stream()))

    def set_clock(self, time):
        """Start the device clock at `time` (the time step follows the
        reference driver's rule, problems/minidisk.cpp:202-206)."""
        self.clock.fill_(0.0)
        self.clock[0] = time

    def clock_time(self):
        """The device clock's time and last time step."""
        t, dt = self.clock.cpu().tolist()
        return t, dt

    def rk2_step(self, u, out, time, dt, rk_b=0.5):
        """One whole RK2 step u -> out in one launch (the fused kernel: the
        first stage's state is never stored); equal to the two stages."""
        check(lib().m3p_minidisk_rk2_step(_ptr(u), _ptr(out), _ptr(self.nbr), _ptr(self.bcoord), self.mesh.nblk,
                                          time, dt, rk_b, ctypes.byref(self.data), self.precision, _stream()))

    def step_fused(self, dt):
        """step(dt) through the fused RK2 kernel (mixed precision)."""
        t = self.time
        self.rk2_step(self.u, self.u2, t, dt)
        self.u, self.u2 = self.u2, self.u
        self.time = t * 0.5 + (t + dt) * (1.0 - 0.5)

    def device_step(self, u, out, rk_order=2, fused=False):
        """One step from u into out on the device clock: the clock kernel,
        then two stages (RK2, or one fused RK2 launch with fused=True) or one
        stage (forward Euler, the reference driver's substep,
        problems/minidisk.cpp:172-186)."""
        check(lib().m3p_minidisk_clock(ctypes.byref(self.data), _ptr(self.clock), _ptr(self.orbit), rk_order,
                                       _stream()))
        if rk_order == 2 and fused:
            check(lib().m3p_minidisk_rk2_step_dev(_ptr(u), _ptr(out), _ptr(self.nbr), _ptr(self.bcoord),
                                                  self.mesh.nblk, _ptr(self.orbit),
                                                  ctypes.c_void_p(self.clock.data_ptr() + 8), 0.5,
                                                  ctypes.byref(self.data), self.precision, _stream()))
        elif rk_order == 2:
            self.stage_dev(u, None, self.u1, 0)
            self.stage_dev(self.u1, u, out, 1)
        else:
            self.stage_dev(u, None, out, 0)

    def device_step_pair(self, rk_order=2, fused=False):
        """Two device-clocked steps that end in the buffer they began in
        (u -> u2 -> u): the unit a Graph captures."""
        self.device_step(self.u, self.u2, rk_order, fused)
        self.device_step(self.u2, self.u, rk_order, fused)

    def autotune(self, reps=5):
        """Pick the fastest compiled tile shape for this mesh (blocking); returns (tj, h, ms)."""
        tj, h, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_float(0)
        check(lib().m3p_autotune_minidisk(_ptr(self.u), _ptr(self.u1), _ptr(self.nbr), _ptr(self.bcoord),
                                          self.mesh.nblk, ctypes.byref(self.data), self.precision, reps,
                                          ctypes.byref(tj), ctypes.byref(h), ctypes.byref(ms), _stream()))
        return tj.value, h.value, ms.value

    def step(self, dt):
        t = self.time
        self.stage(self.u, None, self.u1, t, dt)
        self.stage(self.u1, self.u, self.u2, t + dt, dt)
        self.u, self.u2 = self.u2, self.u
        # solution time combines like the state: s*b + t*(1-b) (src/module_euler.cpp:357)
        self.time = t * 0.5 + (t + dt) * (1.0 - 0.5)


class Euler1DSolver(_StateSolver):
    """euler1d on a chain of blocks with outflow ends (module_euler.cpp:107-161, mara.cpp:123-143)."""
    nc = 5

    def __init__(self, nblk, block_size, gamma=1.4, plm_theta=1.5, precision="fp64",
                 domain_size=1.0, outflow=True, device="cuda"):
        self.mesh = ChainMesh1D(nblk, block_size, outflow)
        self.gamma, self.theta = gamma, plm_theta
        self.precision = PRECISIONS[precision]
        self.dl = domain_size / float(nblk * block_size)
        self._alloc((nblk, 5, block_size), device)
        self.nbr = torch.as_tensor(self.mesh.neighbours(), device=self.device)

    def stage(self, u, u0, out, dt, rk_b=0.5):
        check(lib().m3p_euler1d_stage(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), self.mesh.nblk, self.mesh.n,
                                      dt, self.dl, rk_b, self.theta, self.gamma, self.precision, _stream()))

    def step(self, dt):
        self.stage(self.u, None, self.u1, dt)
        self.stage(self.u1, self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u


class AmrEuler2DSolver(_StateSolver):
    """euler2d on the leaves of a 2:1-balanced quad-tree (module_euler.cpp:263-323
    over amr::extend_block, src/mesh_amr.hpp:300-393).  All leaves advance with
    one dt; each uses its own level's cell spacing.  fused=True: one launch
    per RK stage reads guard zones through the leaf recipes; fused=False: a
    guard-fill launch writes extended primitive blocks, then the update.
    conservative=True (fused only): the flux through every edge between a
    leaf and finer leaves is corrected to the finer leaves' face fluxes, so
    the update conserves across level jumps (the reference does not)."""
    nc = 5

    def __init__(self, tree: QuadTree, block_size, gamma=5.0 / 3.0, plm_theta=1.5, precision="fp64",
                 hllc=False, domain_size=1.0, fused=True, conservative=False, device="cuda"):
        if conservative and not fused:
            raise ValueError("AmrEuler2DSolver: the conservative mode runs the fused stage")
        self.tree = tree
        self.n = int(block_size)
        self.nleaf = len(tree.leaves)
        self.gamma, self.theta, self.hllc = gamma, plm_theta, int(hllc)
        self.precision = PRECISIONS[precision]
        n = self.n
        self._alloc((self.nleaf, 5, n, n), device)
        self.pe = None if fused else torch.zeros((self.nleaf, 5, n + 4, n + 4), dtype=torch.float64,
                                                 device=self.device)
        self.recipes = torch.as_tensor(tree.recipes(), device=self.device)
        self.dl_host = domain_size / float(n) / (2.0 ** tree.levels().astype(np.float64))
        self.dl = torch.as_tensor(self.dl_host, dtype=torch.float64, device=self.device)
        self.dt_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self.flux_work = (torch.zeros((self.nleaf, 4, 5, n), dtype=torch.float64, device=self.device)
                          if conservative else None)

    def stage(self, u, u0, out, dt, rk_b=0.5, dt_dev=None, cfl=None, dt_next=None):
        if self.flux_work is not None:
            if cfl is not None:
                raise ValueError("AmrEuler2DSolver: the conservative mode has no fused CFL stage")
            check(lib().m3p_amr_euler2d_stage_conservative(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.flux_work),
                                                           _ptr(self.recipes), _ptr(self.dl), self.nleaf, self.n, dt,
                                                           _ptr(dt_dev), rk_b, self.theta, self.gamma, self.hllc,
                                                           self.precision, _stream()))
            return
        if cfl is not None:
            check(lib().m3p_amr_euler2d_stage_cfl(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.pe), _ptr(self.recipes),
                                                  _ptr(self.dl), self.nleaf, self.n, dt, _ptr(dt_dev), rk_b,
                                                  self.theta, self.gamma, self.hllc, cfl, _ptr(dt_next),
                                                  self.precision, _stream()))
            return
        check(lib().m3p_amr_euler2d_stage(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.pe), _ptr(self.recipes),
                                          _ptr(self.dl), self.nleaf, self.n, dt, _ptr(dt_dev), rk_b, self.theta,
                                          self.gamma, self.hllc, self.precision, _stream()))

    def step(self, dt):
        self.stage(self.u, None, self.u1, dt)
        self.stage(self.u1, self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u

    def cfl_dt(self, cfl):
        check(lib().m3p_euler_cfl_dt_blocks(_ptr(self.u), self.nleaf, self.n * self.n, _ptr(self.dl), self.gamma,
                                            cfl, _ptr(self.dt_dev), _stream()))
        return self.dt_dev

    def step_cfl(self, cfl):
        dt_dev = self.cfl_dt(cfl)
        self.stage(self.u, None, self.u1, 0.0, dt_dev=dt_dev)
        self.stage(self.u1, self.u, self.u2, 0.0, dt_dev=dt_dev)
        self.u, self.u2 = self.u2, self.u

    def autotune(self, reps=5):
        """Pick the fastest tile shape for this tree (blocking); returns (tj, h, ms)."""
        dt = float(self.cfl_dt(0.1).item())
        return autotune_tiles(_lib.M3P_KIND_AMR_EULER2D, self.precision, self.n,
                              lambda: self.stage(self.u, None, self.u1, dt), reps)

    def conserved_totals_weighted(self):
        """Totals of U * dA over the leaves (host, fsum)."""
        u = self.state()
        import math
        w = (self.dl_host ** 2)[:, None, None, None] * u
        return np.array([math.fsum(w[:, q].ravel()) for q in range(5)])


class Euler1DEnsembleSolver(_StateSolver):
    """Many independent euler1d problems (outflow ends, cell spacing 1/L each)
    advanced together by one persistent work-list launch per RK stage."""
    nc = 5

    def __init__(self, mesh: EnsembleMesh1D, gamma=1.4, plm_theta=1.5, precision="fp64", device="cuda"):
        self.mesh = mesh
        self.gamma, self.theta = gamma, plm_theta
        self.precision = PRECISIONS[precision]
        self._alloc((mesh.nblk, 5, mesh.n), device)
        self.nbr = torch.as_tensor(mesh.neighbours(), device=self.device)
        self.dl = torch.as_tensor(mesh.cell_spacing(), dtype=torch.float64, device=self.device)
        self.dt_dev = torch.zeros(1, dtype=torch.float64, device=self.device)
        self._mutual = self.links_mutual()

    def stage(self, u, u0, out, dt, rk_b=0.5, dt_dev=None, cfl=None, dt_next=None):
        if cfl is not None:
            check(lib().m3p_euler1d_stage_blocks_cfl(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), _ptr(self.dl),
                                                     self.mesh.nblk, self.mesh.n, dt, _ptr(dt_dev), rk_b, self.theta,
                                                     self.gamma, cfl, _ptr(dt_next), self.precision, _stream()))
            return
        check(lib().m3p_euler1d_stage_blocks(_ptr(u), _ptr(u0), _ptr(out), _ptr(self.nbr), _ptr(self.dl),
                                             self.mesh.nblk, self.mesh.n, dt, _ptr(dt_dev), rk_b, self.theta,
                                             self.gamma, self.precision, _stream()))

    def step(self, dt):
        self.stage(self.u, None, self.u1, dt)
        self.stage(self.u1, self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u

    def rk2_step(self, u, out, dt, rk_b=0.5, dt_dev=None, cfl=None, dt_next=None):
        """One whole RK2 step u -> out in one launch (the fused kernel, mixed
        precision: the first stage's state is never stored); equal to the two
        stages.  With cfl and dt_next the CFL time step of `out` is reduced in
        the same launch."""
        if not self._mutual:
            raise ValueError("the fused RK2 step needs mutual prev/next links (a one-block problem has none); "
                             "use step()")
        if cfl is not None:
            check(lib().m3p_euler1d_rk2_step_blocks_cfl(_ptr(u), _ptr(out), _ptr(self.nbr), _ptr(self.dl),
                                                        self.mesh.nblk, self.mesh.n, dt, _ptr(dt_dev), rk_b,
                                                        self.theta, self.gamma, cfl, _ptr(dt_next), self.precision,
                                                        _stream()))
            return
        check(lib().m3p_euler1d_rk2_step_blocks(_ptr(u), _ptr(out), _ptr(self.nbr), _ptr(self.dl), self.mesh.nblk,
                                                self.mesh.n, dt, _ptr(dt_dev), rk_b, self.theta, self.gamma,
                                                self.precision, _stream()))

    def step_fused(self, dt):
        """step(dt) through the fused RK2 kernel (mixed precision)."""
        self.rk2_step(self.u, self.u2, dt)
        self.u, self.u2 = self.u2, self.u

    def links_mutual(self):
        """True when every prev/next link of the chain table is returned by
        the linked block (what the fused RK2 step needs; a one-block problem,
        whose next link is itself and whose prev end is outflow, is not)."""
        nbr = self.mesh.neighbours()
        b = np.arange(nbr.shape[0])
        prev, nxt = nbr[:, 0], nbr[:, 1]
        ok_prev = (prev < 0) | (nbr[np.maximum(prev, 0), 1] == b)
        ok_next = (nxt < 0) | (nbr[np.maximum(nxt, 0), 0] == b)
        return bool(np.all(ok_prev & ok_next))

    def cfl_dt(self, cfl):
        check(lib().m3p_euler_cfl_dt_blocks(_ptr(self.u), self.mesh.nblk, self.mesh.n, _ptr(self.dl), self.gamma,
                                            cfl, _ptr(self.dt_dev), _stream()))
        return self.dt_dev

    def step_cfl(self, cfl):
        dt_dev = self.cfl_dt(cfl)
        self.stage(self.u, None, self.u1, 0.0, dt_dev=dt_dev)
        self.stage(self.u1, self.u, self.u2, 0.0, dt_dev=dt_dev)
        self.u, self.u2 = self.u2, self.u


class Graph:
    """A hipGraph of asynchronous library calls issued on the current stream.

        with Graph() as g:
            solver.step_pair_cfl(cfl)      # any m3p calls
        g.replay()

    Captured calls must take changing time steps from device memory."""

    def __init__(self):
        self.handle = ctypes.c_void_p(0)

    def __enter__(self):
        check(lib().m3p_graph_begin(_stream()))
        return self

    def __exit__(self, kind, value, tb):
        st = lib().m3p_graph_end(_stream(), ctypes.byref(self.handle))
        if kind is None:
            check(st)
        return False

    def replay(self):
        check(lib().m3p_graph_launch(self.handle, _stream()))

    def __del__(self):
        if self.handle and self.handle.value:
            lib().m3p_graph_destroy(self.handle)
            self.handle = ctypes.c_void_p(0)


def cfl_step_pair(solver, cfl):
    """Two CFL-limited RK2 steps that end in the buffer they began in
    (u -> u1 -> u2, then u2 -> u1 -> u): the unit a Graph captures, since a
    graph replays fixed buffer addresses."""
    u, u1, u2 = solver.u, solver.u1, solver.u2
    for a, b in ((u, u2), (u2, u)):
        solver.u = a
        dt = solver.cfl_dt(cfl)
        solver.stage(a, None, u1, 0.0, dt_dev=dt)
        solver.stage(u1, a, b, 0.0, dt_dev=dt)
    solver.u, solver.u1, solver.u2 = u, u1, u2


def cfl_step_pair_fused(solver, cfl):
    """cfl_step_pair with each step's CFL reduction fused into its second
    stage: solver.dt_dev must hold the time step of the current state
    (solver.cfl_dt(cfl) once, before the first call or graph replay), and each
    step leaves the next step's time step there.  Same results as
    cfl_step_pair, without the separate pass over the state."""
    u, u1, u2 = solver.u, solver.u1, solver.u2
    dt = solver.dt_dev
    for a, b in ((u, u2), (u2, u)):
        solver.stage(a, None, u1, 0.0, dt_dev=dt)
        solver.stage(u1, a, b, 0.0, dt_dev=dt, cfl=cfl, dt_next=dt)


def cfl_step_pair_rk2(solver, cfl):
    """cfl_step_pair_fused with each step one fused RK2 launch (an
    Euler2DSolver or Euler1DEnsembleSolver in mixed precision): the launch
    also reduces the next step's CFL time step into solver.dt_dev.  Same
    results as cfl_step_pair."""
    u, u2 = solver.u, solver.u2
    dt = solver.dt_dev
    for a, b in ((u, u2), (u2, u)):
        solver.rk2_step(a, b, 0.0, dt_dev=dt, cfl=cfl, dt_next=dt)


class MhdSolver:
    """Ideal MHD with constrained transport on a periodic N^3 cube:
    mara::advance_mhd_ct (src/scheme_mhd_v1.cpp:265-299) per RK stage, RK2
    with the stage-wise CFL time step of the reference (examples/mhd3d.cpp:186-216)."""

    def __init__(self, N, cfl=0.33, plm_theta=1.5, gamma=5.0 / 3.0, precision="fp64", device="cuda"):
        self.N, self.cfl, self.theta, self.gamma = N, cfl, plm_theta, gamma
        self.precision = PRECISIONS[precision]
        self.dl = 1.0 / N
        self.device = torch.device(device)
        f64 = dict(dtype=torch.float64, device=self.device)
        shapes = [(5, N, N, N), (N + 1, N, N), (N, N + 1, N), (N, N, N + 1)]
        self.state_ = [torch.zeros(s, **f64) for s in shapes]
        self.stage_ = [torch.zeros(s, **f64) for s in shapes]
        self.next_ = [torch.zeros(s, **f64) for s in shapes]
        nbytes = lib().m3p_mhd_workspace_bytes(N, self.precision)
        self.work = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        self.dts = torch.zeros(2, **f64)
        self.divb = torch.zeros(1, **f64)
        self.time = 0.0

    def set_state(self, uc, bf1, bf2, bf3):
        for t, a in zip(self.state_, (uc, bf1, bf2, bf3)):
            t.copy_(torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64))

    def state(self):
        return [t.cpu().numpy() for t in self.state_]

    def advance(self, src, dst, base=None, rk_b=0.5, dt_slot=0):
        b = base if base is not None else [None] * 4
        dt = self.dts[dt_slot:dt_slot + 1]
        check(lib().m3p_mhd_advance_ct(_ptr(src[0]), _ptr(src[1]), _ptr(src[2]), _ptr(src[3]),
                                       _ptr(b[0]), _ptr(b[1]), _ptr(b[2]), _ptr(b[3]), rk_b,
                                       self.N, self.dl, self.cfl, self.theta, self.gamma,
                                       _ptr(dst[0]), _ptr(dst[1]), _ptr(dst[2]), _ptr(dst[3]), _ptr(dt),
                                       _ptr(self.work), self.work.numel(), self.precision, _stream()))

    def step(self):
        """One RK2 step (control::advance_runge_kutta order 2); returns nothing,
        the two stage time steps stay in self.dts (device)."""
        self.advance(self.state_, self.stage_, dt_slot=0)
        self.advance(self.stage_, self.next_, base=self.state_, dt_slot=1)
        self.state_, self.next_ = self.next_, self.state_

    def max_div_b(self):
        s = self.state_
        check(lib().m3p_mhd_max_div_b(_ptr(s[1]), _ptr(s[2]), _ptr(s[3]), self.N, _ptr(self.divb), _stream()))
        return float(self.divb.item())

    def check_errors(self):
        _lib.check_kernel_errors(torch.cuda.current_stream().cuda_stream)
