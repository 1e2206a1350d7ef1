"""Write golden vectors from the reference's own hot-path code.

Runs in the build container only: it calls oracle/_ref/libmara_ref.so, which
oracle/Makefile compiles from the reference sources where they lie
(problems/minidisk_scheme.cpp plus the header-only physics).  The outputs are
small .npz files of inputs and expected outputs; the reference itself never
travels.

    python tests/golden/make_golden.py            # all fixtures
    python tests/golden/make_golden.py mhd_pcm    # only mhd_pcm.npz
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle  # noqa: E402
from oracle import P  # noqa: E402

SEED = 20261202


def plm_cases(rng):
    ys, ths = [], []
    for th in (1.0, 1.5, 2.0):
        for _ in range(300):
            y = rng.normal(size=3)
            k = rng.integers(0, 6)
            if k == 0:
                y[1] = y[0]
            elif k == 1:
                y[2] = y[1]
            elif k == 2:
                y[:] = y[0]
            elif k == 3:
                y[1] = 0.0
                y[0] = -0.0
            ys.append(y)
            ths.append(th)
    ys = np.array(ys)
    ths = np.array(ths)
    out = np.array([oracle.ref().ref_plm(*y, t) for y, t in zip(ys, ths)])
    return dict(y=ys, theta=ths, g=out)


def euler_cases(rng, m=400):
    pl = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 3)), rng.uniform(0.05, 3, m)])
    pr = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 3)), rng.uniform(0.05, 3, m)])
    axis = rng.integers(1, 4, m).astype(np.int32)
    gamma = rng.choice([1.4, 5.0 / 3.0], m)
    f = np.zeros((m, 5))
    flux = np.zeros((m, 5))
    u = np.zeros((m, 5))
    p2 = np.zeros((m, 5))
    for i in range(m):
        a, b, o = np.ascontiguousarray(pl[i]), np.ascontiguousarray(pr[i]), np.zeros(5)
        oracle.ref().ref_euler_riemann_hlle(P(a), P(b), int(axis[i]), gamma[i], P(o))
        f[i] = o
        o = np.zeros(5)
        oracle.ref().ref_euler_flux(P(a), int(axis[i]), gamma[i], P(o))
        flux[i] = o
        o = np.zeros(5)
        oracle.ref().ref_euler_conserved_density(P(a), gamma[i], P(o))
        u[i] = o
        o = np.zeros(5)
        oracle.ref().ref_euler_recover_primitive(P(np.ascontiguousarray(u[i])), gamma[i], P(o))
        p2[i] = o
    return dict(pl=pl, pr=pr, axis=axis, gamma=gamma, hlle=f, flux=flux, cons=u, recovered=p2)


def iso_cases(rng, m=400):
    pl = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 2))])
    pr = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 2))])
    axis = rng.integers(1, 3, m).astype(np.int32)
    cs2 = rng.uniform(1e-3, 2.0, m)
    f = np.zeros((m, 3))
    for i in range(m):
        o = np.zeros(3)
        oracle.ref().ref_iso2d_riemann_hlle(P(np.ascontiguousarray(pl[i])), P(np.ascontiguousarray(pr[i])),
                                            cs2[i], int(axis[i]), P(o))
        f[i] = o
    return dict(pl=pl, pr=pr, axis=axis, cs2=cs2, hlle=f)


def two_body_cases(rng):
    rows, out = [], []
    for q in (1.0, 0.5, 0.1):
        for e in (0.0, 0.3, 0.7):
            for t in (0.0, 0.37, 3.1, 25.0):
                for om in (0.0, 1.0):
                    st = np.zeros(10)
                    oracle.ref().ref_two_body_state(1.0, 1.0, q, e, t, om, P(st))
                    rows.append((q, e, t, om))
                    out.append(st)
    return dict(args=np.array(rows), state=np.array(out))


def minidisk_cases(rng):
    # solver data: {block_size, depth, omega_frame, mach, nu, domain_radius, softening,
    #               eccentricity, mass_ratio, sink_radius, sink_rate, buffer_rate, buffer_scale, cfl}
    n = 16
    sds, ucs, pes, outs, meta = [], [], [], [], []
    configs = [
        (2, 0, 0, 0.0, 1e-3, dict(domain_radius=1.5, softening=0.05, sink_radius=0.05, q=1.0, e=0.0, om=1.0)),
        (2, 1, 2, 0.7, 2e-3, dict(domain_radius=1.5, softening=0.05, sink_radius=0.05, q=1.0, e=0.0, om=1.0)),
        (2, 3, 1, 2.3, 5e-4, dict(domain_radius=2.0, softening=0.02, sink_radius=0.1, q=0.5, e=0.2, om=0.5)),
        (1, 1, 0, 0.1, 1e-3, dict(domain_radius=8.0, softening=0.05, sink_radius=0.05, q=1.0, e=0.0, om=1.0)),
    ]
    for level, bi, bj, time, dt, c in configs:
        sd = np.array([n, level, c["om"], 10.0, 0.0, c["domain_radius"], c["softening"], c["e"], c["q"],
                       c["sink_radius"], 1e3, 1e2, 0.2, 0.3])
        ui = np.zeros((n, n, 3))
        oracle.ref().ref_minidisk_initial_conserved(P(sd), level, bi, bj, P(ui))
        uc = np.ascontiguousarray(ui * np.exp(0.05 * rng.normal(size=(n, n, 1))))
        pe = np.zeros((n + 4, n + 4, 3))
        pe[..., 0] = rng.uniform(0.5, 1.5, (n + 4, n + 4))
        pe[..., 1:] = rng.normal(0, 1, (n + 4, n + 4, 2))
        out = np.zeros((n, n, 3))
        oracle.ref().ref_minidisk_updated_conserved(P(uc), P(pe), P(sd), level, bi, bj, time, dt, P(out))
        sds.append(sd)
        ucs.append(uc)
        pes.append(pe)
        outs.append(out)
        meta.append((level, bi, bj, time, dt))
    return dict(sd=np.array(sds), uc=np.array(ucs), pe=np.array(pes), out=np.array(outs), meta=np.array(meta))


def minidisk_visc_cases(rng):
    """minidisk::updated_conserved with a nonzero kinematic viscosity (the
    viscous stress of godunov_and_viscous_fluxes, problems/minidisk_scheme.cpp:243-287)."""
    n = 16
    sds, ucs, pes, outs, meta = [], [], [], [], []
    for level, bi, bj, time, dt, nu in ((2, 1, 2, 0.4, 1e-3, 1e-3), (1, 0, 1, 1.1, 5e-4, 0.05), (3, 5, 2, 0.0, 1e-3, 0.2)):
        sd = np.array([n, level, 1.0, 10.0, nu, 1.5, 0.05, 0.0, 1.0, 0.05, 1e3, 1e2, 0.2, 0.3])
        ui = np.zeros((n, n, 3))
        oracle.ref().ref_minidisk_initial_conserved(P(sd), level, bi, bj, P(ui))
        uc = np.ascontiguousarray(ui * np.exp(0.05 * rng.normal(size=(n, n, 1))))
        pe = np.zeros((n + 4, n + 4, 3))
        pe[..., 0] = rng.uniform(0.5, 1.5, (n + 4, n + 4))
        pe[..., 1:] = rng.normal(0, 1, (n + 4, n + 4, 2))
        out = np.zeros((n, n, 3))
        oracle.ref().ref_minidisk_updated_conserved(P(uc), P(pe), P(sd), level, bi, bj, time, dt, P(out))
        sds.append(sd)
        ucs.append(uc)
        pes.append(pe)
        outs.append(out)
        meta.append((level, bi, bj, time, dt))
    return dict(sd=np.array(sds), uc=np.array(ucs), pe=np.array(pes), out=np.array(outs), meta=np.array(meta))


def mhd_cases(rng):
    """mhd::riemann_hlle vectors and one advance_mhd_ct call on an 8^3 periodic cube."""
    m = 300
    g = 5.0 / 3.0
    pl = np.column_stack([rng.uniform(0.2, 2, m), rng.normal(0, 1, (m, 3)), rng.uniform(0.2, 2, m), rng.normal(0, 1, (m, 3))])
    pr = np.column_stack([rng.uniform(0.2, 2, m), rng.normal(0, 1, (m, 3)), rng.uniform(0.2, 2, m), rng.normal(0, 1, (m, 3))])
    bl = rng.normal(0, 1, m)
    axis = rng.integers(1, 4, m).astype(np.int32)
    f = np.zeros((m, 5))
    e = np.zeros((m, 3))
    for i in range(m):
        oracle.ref().ref_mhd_riemann_hlle(P(np.ascontiguousarray(pl[i])), P(np.ascontiguousarray(pr[i])), bl[i],
                                          int(axis[i]), g, P(f[i]), P(e[i]))
    from mara3p_amd.problems import orszag_tang_3d
    N = 8
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=3, perturbation=0.05)
    ucl = np.ascontiguousarray(np.moveaxis(uc, 0, -1))
    uo, o1, o2, o3, dt = np.zeros_like(ucl), np.zeros_like(b1), np.zeros_like(b2), np.zeros_like(b3), np.zeros(1)
    st = oracle.ref().ref_mhd_advance_ct(N, P(ucl), P(b1), P(b2), P(b3), 1.0 / N, 0.33, 1.5, P(uo), P(o1), P(o2), P(o3), P(dt))
    assert st == 0
    return dict(pl=pl, pr=pr, bl=bl, axis=axis, flux=f, efield=e, N=N, uc=ucl, bf1=b1, bf2=b2, bf3=b3,
                uc_out=uo, bf1_out=o1, bf2_out=o2, bf3_out=o3, dt=dt)


def mhd_pcm_cases():
    """One advance_mhd_ct call with plm_theta = 0 (first-order face states,
    src/scheme_mhd_v1.cpp:245-247,278 -> godunov_fluxes :119-131) on an 8^3
    periodic cube: the composition of the mhd_scheme_v2 array functions
    (src/scheme_mhd_v2.cpp) computes the same update."""
    from mara3p_amd.problems import orszag_tang_3d
    N = 8
    uc, b1, b2, b3 = orszag_tang_3d(N, seed=5, perturbation=0.05)
    ucl = np.ascontiguousarray(np.moveaxis(uc, 0, -1))
    uo, o1, o2, o3, dt = np.zeros_like(ucl), np.zeros_like(b1), np.zeros_like(b2), np.zeros_like(b3), np.zeros(1)
    st = oracle.ref().ref_mhd_advance_ct(N, P(ucl), P(b1), P(b2), P(b3), 1.0 / N, 0.33, 0.0, P(uo), P(o1), P(o2), P(o3), P(dt))
    assert st == 0
    return dict(N=N, uc=ucl, bf1=b1, bf2=b2, bf3=b3, uc_out=uo, bf1_out=o1, bf2_out=o2, bf3_out=o3, dt=dt)


def srhd_cases(rng, m=300):
    """srhd::recover_primitive and riemann_solver (all four modes) vectors."""
    gam = rng.choice([4.0 / 3.0, 5.0 / 3.0], m)
    pl = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 3)), rng.uniform(0.01, 5, m)])
    pr = np.column_stack([rng.uniform(0.1, 3, m), rng.normal(0, 2, (m, 3)), rng.uniform(0.01, 5, m)])
    axis = rng.integers(1, 4, m).astype(np.int32)
    face = rng.normal(0, 0.3, m)
    f = np.zeros((m, 4, 5))
    st = np.zeros((m, 4), dtype=np.int32)
    u = np.zeros((m, 5))
    p = np.zeros((m, 5))
    for i in range(m):
        a, b = np.ascontiguousarray(pl[i]), np.ascontiguousarray(pr[i])
        for mode in range(4):
            st[i, mode] = oracle.ref().ref_srhd_riemann(P(a), P(b), int(axis[i]), gam[i], mode, face[i], P(f[i, mode]))
        oracle.ref().ref_srhd_conserved_density(P(a), gam[i], P(u[i]))
        oracle.ref().ref_srhd_recover_primitive(P(np.ascontiguousarray(u[i])), gam[i], 0.0, P(p[i]))
    return dict(gamma=gam, pl=pl, pr=pr, axis=axis, face_speed=face, f=f, status=st, cons=u, recovered=p)


def main():
    if not oracle.ref_available():
        raise SystemExit("oracle/_ref/libmara_ref.so is not built (make -C oracle ref)")
    # fixtures with their own inputs can be rewritten alone: make_golden.py mhd_pcm
    only = sys.argv[1:]
    if "mhd_pcm" in only or not only:
        np.savez_compressed(os.path.join(HERE, "mhd_pcm.npz"), **mhd_pcm_cases())
    if only:
        return
    rng = np.random.default_rng(SEED)
    np.savez_compressed(os.path.join(HERE, "plm.npz"), **plm_cases(rng))
    np.savez_compressed(os.path.join(HERE, "euler_pointwise.npz"), **euler_cases(rng))
    np.savez_compressed(os.path.join(HERE, "iso2d_hlle.npz"), **iso_cases(rng))
    np.savez_compressed(os.path.join(HERE, "two_body.npz"), **two_body_cases(rng))
    np.savez_compressed(os.path.join(HERE, "minidisk_block.npz"), **minidisk_cases(rng))
    np.savez_compressed(os.path.join(HERE, "mhd_ct.npz"), **mhd_cases(rng))
    np.savez_compressed(os.path.join(HERE, "srhd_pointwise.npz"), **srhd_cases(rng))
    np.savez_compressed(os.path.join(HERE, "minidisk_visc_block.npz"), **minidisk_visc_cases(rng))
    print("golden vectors written to", HERE)


if __name__ == "__main__":
    main()
