"""Benchmark: zone-updates/s of the Godunov block update on one MI355X.

Workload (default): the locally isothermal binary minidisk -- [-8, 8]^2 as
16 x 16 blocks of 256^2 zones (16.8 M zones), equal-mass circular binary
(a = 1, GM = 1), softening 0.05, Mach 10, sink radius 0.05, Sigma = exp(0.01 z)
seeded log-normal noise, Keplerian velocities, RK2.  The headline value is the
mixed kernel (FP32 reconstruction and fluxes, FP64 conserved state) taking
each RK2 step in one fused launch (equal value for value to its two stage
launches, which are timed beside it, or the headline with --two-stage); the
FP64 drop-in kernel is timed beside it, and the density L1 difference between
the two after --l1-steps steps is reported.

One "step" = one RK2 step (both stages, in one launch or two) over the whole mesh,
state resident in HBM.  A zone-update is one zone through one RK2 step
(the reference's kzps convention, src/app_problem.cpp:84-88).  Before the
W warmup steps, untimed steps run for --settle-seconds (0.5 s) so that the
K timed steps see the GPU at its working clocks.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload minidisk|kh|sedov_amr|ensemble|mhd]

The other workloads are timed the same way for the record: kh = euler2d
Kelvin-Helmholtz 4096^2 with HLLC; sedov_amr = euler2d Sedov blast on a 2:1 quad-tree
of 1408 leaves of 64^2 (levels 3-6); ensemble = 512 independent 1D Riemann
problems of 512-8192 zones; mhd = ideal
MHD with constrained transport, Orszag-Tang at 256^3 (the stage time step
recomputed on the device every stage).  kh and ensemble take their headline
on the fused RK2 step as well (one launch per step), with the two stage
launches timed beside it.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", choices=["minidisk", "kh", "sedov_amr", "ensemble", "mhd"], default="minidisk")
    ap.add_argument("--l1-steps", type=int, default=1000)
    ap.add_argument("--settle-seconds", type=float, default=0.5,
                    help="untimed steps for this long before the warmup steps (GPU clocks settle)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--two-stage", action="store_true",
                    help="minidisk, kh, ensemble: headline on two stage launches per RK2 step instead of the fused step")
    ap.add_argument("--autotune", action="store_true",
                    help="time every tile shape first and keep the fastest (default: the library's tile table)")
    ap.add_argument("--no-autotune", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--no-graph", action="store_true",
                    help="skip the hipGraph replays (kh, sedov_amr, ensemble: CFL-limited steps; minidisk: device-clocked steps)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="also run one plain 1 GiB device copy (for HBM counter calibration under rocprofv3)")
    return ap.parse_args()


# ---------------------------------------------------------------------------
def make_solver(workload, precision):
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import Euler2DSolver, MinidiskSolver
    if workload == "minidisk":
        data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05,
                                     sink_radius=0.05, mach_number=10.0, omega_frame=1.0, mass_ratio=1.0,
                                     eccentricity=0.0)
        mesh, u0 = problems.minidisk(data, seed=1, noise=0.01)
        s = MinidiskSolver(data, precision)
        s.set_state(u0)
        s.zones = mesh.zones()
        dt_of = lambda t: problems.minidisk_dt(data, t)
        return s, u0, dt_of, 3
    if workload == "sedov_amr":
        from mara3p_amd.amr import nested_tree
        from mara3p_amd.solvers import AmrEuler2DSolver
        tree = nested_tree(6, 1.5)
        u0 = problems.sedov_amr(tree, 64)
        s = AmrEuler2DSolver(tree, 64, gamma=5.0 / 3.0, plm_theta=1.5, precision=precision)
        s.set_state(u0)
        dt = float(s.cfl_dt(0.4).item())
        s.zones = len(tree.leaves) * 64 * 64
        return s, u0, (lambda t: dt), 5
    if workload == "ensemble":
        from mara3p_amd.solvers import Euler1DEnsembleSolver
        mesh, u0 = problems.riemann_ensemble(count=512, block_size=64, seed=1)
        s = Euler1DEnsembleSolver(mesh, gamma=1.4, plm_theta=1.5, precision=precision)
        s.set_state(u0)
        dt = float(s.cfl_dt(0.4).item())
        s.zones = mesh.zones()
        return s, u0, (lambda t: dt), 5
    mesh, u0 = problems.kelvin_helmholtz(depth=4, block_size=256, seed=1)
    s = Euler2DSolver(4, 256, gamma=5.0 / 3.0, plm_theta=1.5, precision=precision, hllc=True)
    s.set_state(u0)
    dt = float(s.cfl_dt(0.4).item())
    s.zones = mesh.zones()
    return s, u0, (lambda t: dt), 5


TILE_KIND = {"kh": 0, "minidisk": 1, "sedov_amr": 2}

# conserved variables (the reference's component order) and the number of
# steps of the mixed-vs-FP64 comparison per workload (minidisk: --l1-steps)
CONSERVED = {"minidisk": ["Sigma", "Sx", "Sy"], "kh": ["D", "S1", "S2", "S3", "E"],
             "sedov_amr": ["D", "S1", "S2", "S3", "E"], "ensemble": ["D", "S1", "S2", "S3", "E"]}
L1_STEPS = {"kh": 200, "sedov_amr": 100, "ensemble": 100}


def l1_report(workload, steps, fused=False):
    """Advance the FP64 and the mixed solver through the same time steps and
    report, per conserved variable, mean |mixed - FP64| (absolute and relative
    to mean |FP64|); periodic workloads also report the total mass of each
    relative to the initial mass.  fused: the mixed solver takes the fused
    RK2 step (equal value for value to its two stage launches)."""
    a, ua, dtf, _ = make_solver(workload, "fp64")
    b, _, _, _ = make_solver(workload, "mixed")
    m0 = math.fsum(ua[:, 0].ravel())
    bstep = b.step_fused if fused else b.step
    for _ in range(steps):
        dt = dtf(getattr(a, "time", 1.0))
        a.step(dt)
        bstep(dt)
    a.check_errors()
    b.check_errors()
    ra, rb = a.state(), b.state()
    per = {}
    for q, name in enumerate(CONSERVED[workload]):
        d = float(np.mean(np.abs(ra[:, q] - rb[:, q])))
        per[name] = {"l1": d, "l1_rel": d / max(float(np.mean(np.abs(ra[:, q]))), 1e-300)}
    rep = {"steps": steps, "per_variable": per}
    if workload in ("minidisk", "kh"):
        rep["mass_fp64"] = math.fsum(ra[:, 0].ravel()) / m0
        rep["mass_mixed"] = math.fsum(rb[:, 0].ravel()) / m0
    del a, b
    if workload == "minidisk":
        rep.update(minidisk_mass_drift(steps, fused))
        rep["mixed_step"] = "fused RK2 step" if fused else "two stage launches"
    return rep


def minidisk_mass_drift(steps, fused=False):
    """Relative drift of the total mass after `steps` RK2 steps of the minidisk
    mesh with the sink and buffer rates at zero (the only terms that add or
    remove mass), in each kernel: |M(t) - M(0)| / M(0)."""
    from mara3p_amd import MinidiskData, problems
    from mara3p_amd.solvers import MinidiskSolver
    data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05,
                                 sink_radius=0.05, mach_number=10.0, omega_frame=1.0, mass_ratio=1.0,
                                 eccentricity=0.0, sink_rate=0.0, buffer_rate=0.0)
    _, u0 = problems.minidisk(data, seed=1, noise=0.01)
    m0 = math.fsum(u0[:, 0].ravel())
    out = {"mass_drift_steps": steps, "mass_drift_note": "sink and buffer rates 0; |M(t) - M(0)| / M(0)"}
    for precision in ("fp64", "mixed"):
        s = MinidiskSolver(data, precision)
        s.set_state(u0)
        step = s.step_fused if (fused and precision == "mixed") else s.step
        for _ in range(steps):
            step(problems.minidisk_dt(data, s.time))
        s.check_errors()
        out["mass_drift_" + precision] = abs(math.fsum(s.state()[:, 0].ravel()) - m0) / m0
        del s
    return out


def tile_in_use(workload, precision):
    """The stage kernel's tile shape (columns, rows) from the library's table."""
    from mara3p_amd._lib import lib, PRECISIONS
    if workload not in TILE_KIND:
        return None
    tj, h = ctypes.c_int(0), ctypes.c_int(0)
    lib().m3p_get_stage2d_tile(TILE_KIND[workload], PRECISIONS[precision], ctypes.byref(tj), ctypes.byref(h))
    return [tj.value, h.value]


def settle(step, seconds, count=None):
    """Untimed steps for `seconds` of wall time (or exactly `count` steps),
    before the warmup steps: a timed region of a few milliseconds that starts
    after an idle stretch (the solver's set-up on the host) measured its
    kernels up to 10 % slower than the same kernels at working clocks.
    Returns the number of steps taken."""
    import torch
    n, t0 = 0, time.perf_counter()
    while (n < count) if count is not None else (time.perf_counter() - t0 < seconds):
        step()
        n += 1
        if n % 8 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    return n


def timed_run(s, dt_of, steps, warmup, autotune=False, settle_seconds=0.0):
    """Time `steps` RK2 steps; per-launch durations from HIP events on the launch stream."""
    import torch
    if autotune and hasattr(s, "autotune"):
        s.autotune()
    settle(lambda: s.step(dt_of(getattr(s, "time", 1.0))), settle_seconds)
    for _ in range(warmup):
        s.step(dt_of(getattr(s, "time", 1.0)))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(2 * steps)]
    orig_stage = s.stage
    counter = [0]

    def stage(*a, **k):
        e0, e1 = ev[counter[0]]
        e0.record(stream)
        orig_stage(*a, **k)
        e1.record(stream)
        counter[0] += 1
    s.stage = stage
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step(dt_of(getattr(s, "time", 1.0)))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s.stage = orig_stage
    s.check_errors()
    durs = [a.elapsed_time(b) for a, b in ev]      # ms
    return wall, float(np.mean(durs[0::2])), float(np.mean(durs[1::2]))


# ---------------------------------------------------------------------------
def graph_run(s, steps, warmup, cfl=0.4, rk2=False):
    """The stages captured in a hipGraph: one replay is two CFL-limited RK2
    steps.  Fused: each step's second stage also reduces the next step's CFL
    time step (the library's *_stage_cfl entry points); separate: a CFL
    reduction pass before each step; rk2 (kh, ensemble): each step one fused
    RK2 launch that also reduces the next step's time step.  Whole-step wall
    time over the replays."""
    import torch
    from mara3p_amd.solvers import Graph, cfl_step_pair, cfl_step_pair_fused, cfl_step_pair_rk2
    reps = max(1, steps // 2)
    out = {}
    pairs = [("fused", cfl_step_pair_fused), ("separate", cfl_step_pair)]
    if rk2:
        pairs.append(("rk2", cfl_step_pair_rk2))
    for name, pair in pairs:
        torch.cuda.synchronize()
        stream = torch.cuda.Stream()       # capture needs a stream other than the default one
        with torch.cuda.stream(stream):
            s.cfl_dt(cfl)
            with Graph() as g:
                pair(s, cfl)
            for _ in range(max(1, warmup // 2)):
                g.replay()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                g.replay()
            stream.synchronize()
            wall = time.perf_counter() - t0
            s.check_errors()
        torch.cuda.synchronize()
        del g
        out[name] = wall
    res = {"value": s.zones * 2 * reps / out["fused"], "ms_per_step": 1e3 * out["fused"] / (2 * reps),
           "separate_cfl_ms_per_step": 1e3 * out["separate"] / (2 * reps), "steps": 2 * reps,
           "note": "hipGraph replays of two CFL-limited RK2 steps; value and ms_per_step with the CFL reduction "
                   "fused into each step's second stage, separate_cfl_ms_per_step with a reduction pass per step"}
    if rk2:
        res["rk2_fused_ms_per_step"] = 1e3 * out["rk2"] / (2 * reps)
        res["note"] += ", rk2_fused_ms_per_step with each step one fused RK2 launch reducing the next time step"
    return res


def fused_run(s, dt_of, steps, warmup, settle_seconds):
    """K RK2 steps through the fused kernel (one launch per step, the first
    stage's state never stored): wall time and the mean launch duration from
    HIP events on the launch stream."""
    import torch
    t_of = lambda: getattr(s, "time", 0.0)
    settle(lambda: s.step_fused(dt_of(t_of())), settle_seconds)
    for _ in range(warmup):
        s.step_fused(dt_of(t_of()))
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    orig = s.rk2_step
    counter = [0]

    def rk2_step(*a, **k):
        e0, e1 = ev[counter[0]]
        e0.record(stream)
        orig(*a, **k)
        e1.record(stream)
        counter[0] += 1
    s.rk2_step = rk2_step
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.step_fused(dt_of(t_of()))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    s.rk2_step = orig
    s.check_errors()
    return wall, float(np.mean([a.elapsed_time(b) for a, b in ev]))


def graph_run_minidisk(s, steps, warmup, settle_seconds):
    """The minidisk steps on the device clock (m3p_minidisk_clock: dt from the
    reference driver's rule and the binary's state computed on the device)
    captured in a hipGraph: one replay is two RK2 steps (a clock kernel and
    two stage launches each), or two forward-Euler steps, the reference
    driver's substep (a clock kernel and one stage launch each).  Whole-step
    wall time over the replays."""
    import torch
    from mara3p_amd.solvers import Graph
    reps = max(1, steps // 2)
    out = {}
    for order, fused in ((2, False), (1, False), (2, True)):
        torch.cuda.synchronize()
        stream = torch.cuda.Stream()       # capture needs a stream other than the default one
        with torch.cuda.stream(stream):
            s.set_clock(getattr(s, "time", 0.0))
            with Graph() as g:
                s.device_step_pair(order, fused)
            settle(g.replay, settle_seconds)
            for _ in range(max(1, warmup // 2)):
                g.replay()
            stream.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                g.replay()
            stream.synchronize()
            wall = time.perf_counter() - t0
            s.check_errors()
        torch.cuda.synchronize()
        del g
        out[(order, fused)] = wall
    return {"value": s.zones * 2 * reps / out[(2, False)], "ms_per_step": 1e3 * out[(2, False)] / (2 * reps),
            "steps": 2 * reps, "fused_ms_per_step": 1e3 * out[(2, True)] / (2 * reps),
            "forward_euler_ms_per_step": 1e3 * out[(1, False)] / (2 * reps),
            "note": "hipGraph replays of two device-clocked RK2 steps (value, ms_per_step) or two forward-Euler "
                    "steps (the reference driver's substep, one stage launch each); the clock kernel computes "
                    "dt and the binary's state on the device"}


def host_cpu():
    """The host CPU model and logical CPU count (from /proc/cpuinfo)."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "nproc": os.cpu_count()}


def cpu_baseline(workload, u0, dt, min_seconds=3.0):
    """Reference CPU path on the host cores, timed on a bounded sample.

    minidisk: the reference's own functions compiled from its sources
    (oracle/_ref, -O3, x86-64-v3), the mesh held in the reference's array
    types, one RK2 step = per block minidisk::recover_primitive_array, then
    per block the guard fill and minidisk::updated_conserved, twice, and the
    weighted sum, every per-block task a job of the reference's
    mara::ThreadPool (src/parallel_thread_pool.hpp, the scheduling of
    src/parallel_computable.cpp:248-296) with one worker per host core; the
    calling thread only waits.  One RK2 step of warm-up, then whole RK2 steps
    until at least `min_seconds` have passed.  Without the reference build
    (and for kh) the oracle's C restatement with OpenMP over blocks."""
    import oracle
    from oracle import OrcMinidisk, P, orc
    from mara3p_amd import MinidiskData
    from mara3p_amd.mesh import to_reference_layout
    threads = max(1, min(os.cpu_count() or 1, 16))
    u = np.ascontiguousarray(to_reference_layout(u0))
    nblk = u.shape[0]
    zones = nblk * u.shape[1] * u.shape[2]

    def timed(step):
        step()                               # warm-up
        n, t0 = 0, time.perf_counter()
        while True:
            step()
            n += 1
            el = time.perf_counter() - t0
            if el >= min_seconds:
                return n, el

    if workload != "minidisk":
        state = u.copy()
        steps, el = timed(lambda: orc().orc_euler2d_mesh_rk2(P(state), 16, 256, dt, 1.0 / 4096, 1.5, 5.0 / 3.0,
                                                            1, 1, threads))
        return {"value": zones * steps / el, "unit": "zone-updates/s", "cores": threads, "kind": "port",
                "sample": "%d RK2 steps of the full 4096^2 mesh after one of warm-up (%.1f s): oracle C "
                          "restatement (-O2), OpenMP over blocks" % (steps, el)}
    data = MinidiskData.defaults(block_size=256, depth=4, domain_radius=8.0, softening_length=0.05, sink_radius=0.05)
    sd = OrcMinidisk.from_data(data)
    if oracle.ref_available():
        h = oracle.ref().ref_minidisk_mesh_create(P(u), P(sd.as_array()), threads)
        try:
            t = [0.0]

            def step():
                oracle.ref().ref_minidisk_mesh_rk2(h, t[0], dt, 1)
                t[0] += dt
            steps, el = timed(step)
        finally:
            oracle.ref().ref_minidisk_mesh_destroy(h)
        return {"value": zones * steps / el, "unit": "zone-updates/s", "cores": threads, "kind": "reference",
                "sample": "%d RK2 steps of the full 16x16x256^2 mesh after one of warm-up (%.1f s): the reference's "
                          "recover_primitive_array, guard fill and minidisk::updated_conserved compiled from its "
                          "sources (-O3, x86-64-v3), one mara::ThreadPool task per block per pass, %d workers, "
                          "calling thread waiting" % (steps, el, threads)}
    state = u.copy()
    t = [0.0]

    def step():
        orc().orc_minidisk_mesh_rk2(P(state), sd, t[0], dt, 1, threads)
        t[0] += dt
    steps, el = timed(step)
    return {"value": zones * steps / el, "unit": "zone-updates/s", "cores": threads, "kind": "port",
            "sample": "%d RK2 steps of the full 16x16x256^2 mesh after one of warm-up (%.1f s): oracle C "
                      "restatement (-O2), OpenMP over blocks" % (steps, el)}


def traffic_from_profiles(workload, precision):
    """HBM bytes per launch from committed rocprofv3 PMC summaries, if present."""
    path = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (workload, precision))
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("hbm_bytes_per_launch")


def bench_mhd(args):
    """3D MHD CT at 256^3: K RK2 steps, per-launch times from HIP events
    around each advance_mhd_ct stage (several kernels per stage)."""
    import torch
    from mara3p_amd import problems
    from mara3p_amd.solvers import MhdSolver
    N = 256
    uc, b1, b2, b3 = problems.orszag_tang_3d(N, seed=1)
    out, final = {}, {}
    settled = None
    for precision in ("mixed", "fp64"):
        s = MhdSolver(N, precision=precision)
        s.set_state(uc, b1, b2, b3)
        # the FP64 solver takes as many settling steps as the mixed one, so
        # that both end after the same number of steps for the L1 comparison
        settled = settle(s.step, args.settle_seconds, settled)
        for _ in range(args.warmup):
            s.step()
        torch.cuda.synchronize()
        stream = torch.cuda.current_stream()
        ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
        t0 = time.perf_counter()
        for e0, e1 in ev:
            e0.record(stream)
            s.step()
            e1.record(stream)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        s.check_errors()
        step_ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
        zones = N ** 3
        # RK2: stage 1 reads 8 fields and writes 8, stage 2 also reads the base: 5 x 8 x 8 B per zone-update
        alg = 5 * 8 * 8 * zones
        out[precision] = {"value": zones * args.steps / wall, "ms_per_step": 1e3 * wall / args.steps,
                          "event_ms_per_step": step_ms, "hbm_frac": alg / (step_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                          "max_div_b": s.max_div_b()}
        if args.l1_steps > 0:
            final[precision] = s.state()
        del s
        torch.cuda.synchronize()
    m = out["mixed"]
    result = {"metric": "zone-updates/s (3D MHD CT)", "value": m["value"], "unit": "zone-updates/s", "n_gpus": 1,
              "steps": args.steps, "warmup": args.warmup, "settle_s": args.settle_seconds,
              "ms_per_step": m["ms_per_step"], "higher_is_better": True,
              "scaling": "weak", "vs_baseline": None, "dtype": "f32-flux/f64-state", "data": "synthetic",
              "config": {"workload": "ideal MHD CT, Orszag-Tang 256^3 periodic, PLM+HLLE, RK2", "zones": N ** 3},
              "roofline": {"bound": "hbm", "achieved": m["hbm_frac"] * HBM_PEAK_GBS, "peak": HBM_PEAK_GBS,
                           "unit": "GB/s", "frac": m["hbm_frac"], "traffic": traffic_from_profiles("mhd", "mixed"),
                           "note": "per RK2 step over all kernels of both stages (achieved and traffic)"},
              "fp64": out["fp64"], "max_div_b_mixed": m["max_div_b"]}
    if final:
        # both precisions took settle + warmup + steps RK2 steps, each with its own CFL time step per stage
        names = ["D", "S1", "S2", "S3", "E", "B1", "B2", "B3"]
        fa, fb = final["fp64"], final["mixed"]
        arrays = [(fa[0][q], fb[0][q]) for q in range(5)] + [(fa[1 + a], fb[1 + a]) for a in range(3)]
        per = {}
        for name, (x, y) in zip(names, arrays):
            d = float(np.mean(np.abs(x - y)))
            per[name] = {"l1": d, "l1_rel": d / max(float(np.mean(np.abs(x))), 1e-300)}
        result["l1_mixed_vs_fp64"] = {"steps": settled + args.warmup + args.steps, "per_variable": per}
    print(json.dumps(result), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: measures one GPU only (the library has no multi-GPU path); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.set_device(0)
    if args.copy_calibration:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")   # 1 GiB
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if args.workload == "mhd":
        return bench_mhd(args)
    # The numerics comparison runs first: besides its own figures, its
    # thousands of stage launches bring the GPU to its working clocks before
    # the timed steps (a short timed region right after start-up measured
    # about 10 % slow).
    fused = args.workload in ("minidisk", "kh", "ensemble") and not args.two_stage
    rep = None
    if args.l1_steps > 0:
        steps = args.l1_steps if args.workload == "minidisk" else min(args.l1_steps, L1_STEPS[args.workload])
        rep = l1_report(args.workload, steps, fused)
    s, u0, dt_of, nc = make_solver(args.workload, "mixed")
    zones = s.zones
    wall, d1, d2 = timed_run(s, dt_of, args.steps, args.warmup, args.autotune, args.settle_seconds)
    tile = tile_in_use(args.workload, "mixed")
    graph = None
    if not args.no_graph:
        if args.workload in ("kh", "sedov_amr", "ensemble"):
            graph = graph_run(s, args.steps, args.warmup, rk2=args.workload in ("kh", "ensemble"))
        elif args.workload == "minidisk":
            graph = graph_run_minidisk(s, args.steps, args.warmup, args.settle_seconds)
    ms_per_step = 1e3 * wall / args.steps
    value = zones * args.steps / wall
    # algorithmic bytes: stage 1 reads U0 writes U1 (2 nc 8 B/zone), stage 2 reads U1, U0 writes U (3 nc 8 B/zone)
    bytes_per_launch = 0.5 * (2 + 3) * nc * 8 * zones
    kern_ms = 0.5 * (d1 + d2)
    achieved = bytes_per_launch / (kern_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles(args.workload, "mixed")
    kernel, launch_ms = "stage2d_kernel (two launches per RK2 step)", kern_ms
    two_stage = None
    if fused:
        # the headline: one fused launch per RK2 step, the two stage launches
        # (equal value for value) reported beside it
        two_stage = {"value": value, "ms_per_step": ms_per_step, "stage_ms": {"stage1": d1, "stage2": d2},
                     "tile": tile_in_use(args.workload, "mixed"),
                     "roofline": {"achieved": achieved, "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}}
        fw, fms = fused_run(s, dt_of, args.steps, args.warmup, args.settle_seconds)
        value = zones * args.steps / fw
        ms_per_step = 1e3 * fw / args.steps
        achieved = 2 * bytes_per_launch / (fms * 1e-3) / 1e9
        traffic = traffic_from_profiles(args.workload, "fused")
        kernel, launch_ms = "rk2_fused_kernel (one launch per RK2 step)", fms
    result = {
        "metric": "zone-updates/s, FP32-mixed vs FP64 GPU at matched L1 error; % of HBM peak",
        "value": value,
        "unit": "zone-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "settle_s": args.settle_seconds,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-flux/f64-state",
        "data": "synthetic",
        "config": {},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": kernel, "launch_ms": launch_ms,
                     "traffic_frac": (traffic / (launch_ms * 1e-3) / 1e9 / HBM_PEAK_GBS) if traffic else None},
        "stage_ms": {"stage1": d1, "stage2": d2},
        "tile": tile,
    }
    if fused:
        del result["stage_ms"], result["tile"]
        result["roofline"]["note"] = ("achieved: the RK2 step's algorithmic bytes (5 nc 8 B per zone: U0, U1 "
                                      "and U1, U0, U of the two stages) over the fused launch's duration, which "
                                      "can exceed the peak since the launch never stores U1; traffic: the HBM "
                                      "bytes the fused launch moves; traffic_frac: those over the launch's "
                                      "duration, as a fraction of the peak")
        result["two_stage"] = two_stage
    if graph:
        result["graph"] = graph
    if args.workload == "minidisk":
        result["config"] = {"workload": "minidisk iso2d+two_body, 16x16 blocks of 256^2 (16.8M zones), RK2, mixed"
                                        + (", one fused launch per step" if fused else ", two stage launches per step"),
                            "zones": zones, "blocks": 256, "block_size": 256}
    elif args.workload == "kh":
        result["config"] = {"workload": "euler2d Kelvin-Helmholtz 4096^2 periodic (16x16 blocks of 256^2), PLM+HLLC, RK2, "
                                        "CFL 0.4, mixed"
                                        + (", one fused launch per step" if fused else ", two stage launches per step"),
                            "zones": zones, "blocks": 256, "block_size": 256}
    elif args.workload == "sedov_amr":
        result["config"] = {"workload": "euler2d Sedov blast on a 2:1 quad-tree, 1408 leaves of 64^2 (levels 3-6), "
                                        "guard fill across level jumps, PLM+HLLE, RK2", "zones": zones,
                            "blocks": zones // 4096, "block_size": 64}
    else:
        result["config"] = {"workload": "512 independent 1D Riemann problems of 512-8192 zones in blocks of 64, "
                                        "PLM+HLLE, RK2, mixed"
                                        + (", one fused launch per step" if fused
                                           else ", one work-list launch per stage"), "zones": zones,
                            "blocks": zones // 64, "block_size": 64}

    if not args.no_fp64:
        s64, _, _, _ = make_solver(args.workload, "fp64")
        k64 = max(10, args.steps // 4)
        wall64, e1, e2 = timed_run(s64, dt_of, k64, max(2, args.warmup // 4), args.autotune, args.settle_seconds)
        tile64 = tile_in_use(args.workload, "fp64")
        v64 = zones * k64 / wall64
        result["fp64"] = {"value": v64, "ms_per_step": 1e3 * wall64 / k64, "stage_ms": {"stage1": e1, "stage2": e2},
                          "tile": tile64,
                          "hbm_frac": bytes_per_launch / (0.5 * (e1 + e2) * 1e-3) / 1e9 / HBM_PEAK_GBS}
        result["mixed_over_fp64"] = value / v64
        del s64

    if rep is not None:
        result["l1_mixed_vs_fp64"] = rep
        if args.workload == "minidisk":
            result["l1_rho_mixed_vs_fp64"] = {"l1": rep["per_variable"]["Sigma"]["l1"],
                                              "mass_fp64": rep["mass_fp64"], "mass_mixed": rep["mass_mixed"]}
            result["l1_rho_mixed_vs_fp64"]["steps"] = rep["steps"]
            result["mass_drift_fp64"] = rep["mass_drift_fp64"]
            result["mass_drift_mixed"] = rep["mass_drift_mixed"]

    if not args.no_cpu_baseline and args.workload in ("minidisk", "kh"):
        try:
            result["cpu_baseline"] = cpu_baseline(args.workload, u0, dt_of(1.0))
            result["cpu_baseline"]["host"] = host_cpu()
        except Exception as e:  # the baseline is reported, never required
            result["cpu_baseline"] = {"value": None, "error": str(e)}
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
