"""CPU-side checks of the host layer: the C ABI library loads and exports every
symbol include/mara3p_amd.h declares (no compute calls), mesh tables, layout
conversions and the seeded initial conditions."""
import ctypes
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mara3p_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(m3p_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import mara3p_amd
    handle = ctypes.CDLL(mara3p_amd.LIBRARY_PATH)
    syms = declared_symbols()
    assert len(syms) >= 12
    for s in syms:
        assert hasattr(handle, s), s
    # and the Python binding covers exactly the declared ABI
    from mara3p_amd._lib import SIGNATURES
    assert sorted(SIGNATURES) == syms


def test_version_and_error_text():
    import mara3p_amd
    assert mara3p_amd.lib().m3p_version() >= 1
    assert isinstance(mara3p_amd._lib.last_error(), str)


def test_uniform_mesh_neighbours_wrap():
    from mara3p_amd.mesh import UniformMesh2D
    m = UniformMesh2D(2, 16)
    nbr = m.neighbours()
    ij = m.block_coordinates()
    for b in range(m.nblk):
        i, j = ij[b]
        assert tuple(ij[nbr[b, 0]]) == ((i - 1) % 4, j)
        assert tuple(ij[nbr[b, 1]]) == ((i + 1) % 4, j)
        assert tuple(ij[nbr[b, 2]]) == (i, (j - 1) % 4)
        assert tuple(ij[nbr[b, 3]]) == (i, (j + 1) % 4)


def test_chain_mesh_outflow_single_block_quirk():
    from mara3p_amd.mesh import ChainMesh1D
    assert ChainMesh1D(1, 16).neighbours().tolist() == [[-1, 0]]
    assert ChainMesh1D(3, 16).neighbours().tolist() == [[-1, 1], [0, 2], [1, -1]]
    assert ChainMesh1D(3, 16, outflow=False).neighbours().tolist() == [[2, 1], [0, 2], [1, 0]]


def test_layout_round_trip():
    from mara3p_amd.mesh import to_device_layout, to_reference_layout
    a = np.random.default_rng(0).random((4, 8, 8, 5))
    d = to_device_layout(a)
    assert d.shape == (4, 5, 8, 8)
    assert np.array_equal(d[1, 3, 2, 7], a[1, 2, 7, 3])
    assert np.array_equal(to_reference_layout(d), a)


def test_initial_conditions_are_decomposition_independent():
    from mara3p_amd import problems
    from mara3p_amd.mesh import UniformMesh2D
    m1, u1 = problems.kelvin_helmholtz(depth=1, block_size=32, seed=7)
    m2, u2 = problems.kelvin_helmholtz(depth=2, block_size=16, seed=7)

    def glob(m, u):
        return u.reshape(m.nb, m.nb, 5, m.n, m.n).transpose(2, 0, 3, 1, 4).reshape(5, m.nb * m.n, m.nb * m.n)
    assert np.array_equal(glob(m1, u1), glob(m2, u2))


def test_minidisk_ic_matches_oracle_without_noise():
    import oracle
    from oracle import OrcMinidisk, P
    from mara3p_amd import MinidiskData, problems
    data = MinidiskData.defaults(block_size=16, depth=2, domain_radius=8.0, softening_length=0.05)
    mesh, u = problems.minidisk(data, noise=0.0)
    sd = OrcMinidisk.from_data(data)
    for b in (0, 5, 15):
        ref = np.zeros((16, 16, 3))
        i, j = mesh.block_coordinates()[b]
        oracle.orc().orc_minidisk_initial_conserved(ctypes.byref(sd), 2, int(i), int(j), P(ref))
        assert np.allclose(np.moveaxis(u[b], 0, -1), ref, rtol=1e-14, atol=1e-15)


def test_lognormal_noise_statistics():
    from mara3p_amd import problems
    z = problems.normal(3, np.arange(200000, dtype=np.uint64))
    assert abs(z.mean()) < 0.01 and abs(z.std() - 1) < 0.01
