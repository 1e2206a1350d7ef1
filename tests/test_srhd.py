"""Special-relativistic hydrodynamics (src/physics_srhd.hpp): the oracle against
the reference's golden vectors and known-answer tests, then the HIP kernels
(FP64, exact) against the oracle."""
import ctypes
import os

import numpy as np
import pytest

from oracle import P, orc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def golden():
    return np.load(os.path.join(GOLDEN, "srhd_pointwise.npz"))


def test_srhd_golden_vectors():
    g = golden()
    for i in range(len(g["pl"])):
        for mode in range(4):
            f = np.zeros(5)
            st = orc().orc_srhd_riemann(P(c(g["pl"][i])), P(c(g["pr"][i])), int(g["axis"][i]), float(g["gamma"][i]),
                                        mode, float(g["face_speed"][i]), P(f))
            assert st == g["status"][i, mode]
            assert np.array_equal(f, g["f"][i, mode])
        u, p = np.zeros(5), np.zeros(5)
        orc().orc_srhd_conserved_density(P(c(g["pl"][i])), float(g["gamma"][i]), P(u))
        assert np.array_equal(u, g["cons"][i])
        assert orc().orc_srhd_recover_primitive(P(u), float(g["gamma"][i]), 0.0, P(p)) == 0
        assert np.array_equal(p, g["recovered"][i])


@pytest.mark.parametrize("p0", [(1.0, 0, 0, 0, 1.0), (1.5, 0, 0, 0, 1e3), (1.0, 0, 0, 0, 1.0), (5.7, 1.0, 2.0, 3.0, 5.0)])
def test_srhd_round_trip_kat(p0):
    """test_srhd, src/physics_srhd.hpp:526-536: cons -> prim at 1e-10, gamma 4/3."""
    p0 = c(p0)
    u, p1 = np.zeros(5), np.zeros(5)
    orc().orc_srhd_conserved_density(P(p0), 4.0 / 3.0, P(u))
    assert orc().orc_srhd_recover_primitive(P(u), 4.0 / 3.0, 0.0, P(p1)) == 0
    assert np.all(np.abs(p1 - p0) < 1e-10)


@pytest.mark.parametrize("axis", [1, 2, 3])
def test_srhd_contact_speed_kat(axis):
    """test_srhd :538-549: the HLLC contact speed of a pure density jump equals the
    flow velocity, and the mass flux through a face moving with it is < 1e-12."""
    u = [0.5 * (axis == 1), 0.6 * (axis == 2), 0.7 * (axis == 3)]
    pl, pr = c([1.0, *u, 10.0]), c([4.0, *u, 10.0])
    f = np.zeros(5)
    assert orc().orc_srhd_riemann(P(pl), P(pr), axis, 4.0 / 3.0, 2, 0.0, P(f)) == 0
    cs = f[0]
    W = np.sqrt(1 + np.sum(np.square(u)))
    assert abs(cs - u[axis - 1] / W) < 1e-12
    orc().orc_srhd_riemann(P(pl), P(pr), axis, 4.0 / 3.0, 3, cs, P(f))
    assert abs(f[0]) < 1e-12


torch = pytest.importorskip("torch")


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")
def test_gpu_srhd_matches_oracle_exactly():
    import mara3p_amd
    g = golden()
    m = len(g["pl"])
    ptr = lambda t: ctypes.c_void_p(t.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    dev = lambda a: torch.as_tensor(c(a), device="cuda")
    for gam in (4.0 / 3.0, 5.0 / 3.0):
        sel = np.isclose(g["gamma"], gam)
        pl, pr, ax = g["pl"][sel], g["pr"][sel], g["axis"][sel]
        for axis in (1, 2, 3):
            k = ax == axis
            n = int(k.sum())
            if n == 0:
                continue
            dpl, dpr = dev(pl[k].T), dev(pr[k].T)
            for mode in (0, 1):
                out = torch.zeros((5, n), dtype=torch.float64, device="cuda")
                mara3p_amd.check(mara3p_amd.lib().m3p_srhd_riemann_array(ptr(dpl), ptr(dpr), ptr(out), n, axis, gam,
                                                                         mode, 0.0, stream))
                ref = g["f"][sel][k][:, mode]
                assert np.array_equal(out.cpu().numpy().T, ref)
        u = dev(g["cons"][sel].T)
        p = torch.zeros_like(u)
        mara3p_amd.check(mara3p_amd.lib().m3p_srhd_recover_primitive_array(ptr(u), ptr(p), u.shape[1], gam, 0.0, stream))
        assert np.array_equal(p.cpu().numpy().T, g["recovered"][sel])
