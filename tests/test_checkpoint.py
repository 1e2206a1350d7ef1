"""Checkpoints: HDF5 files in the reference's layout (src/app_problem.cpp:55-78,
src/module_euler_io.cpp:81-102, block names by mesh::format_block_index,
src/mesh_block_index.hpp:353-408), written by this package's own HDF5 writer
and read back by its reader and, where the image has it, by the HDF5
library's h5dump."""
import os
import subprocess

import numpy as np
import pytest

from mara3p_amd import checkpoint, hdf5


@pytest.mark.parametrize("index,text", [((0, 0, 0), "0:0-0"), ((3, 5, 7), "3:5-7"), ((4, 3, 12), "4:03-12"),
                                        ((7, 100, 9), "7:100-009"), ((10, 5, 1023), "10:0005-1023"),
                                        ((3, 5), "3:5"), ((6, 42), "6:42")])
def test_format_block_index(index, text):
    """Zero padding to 1 + log10(2^level) digits, truncated (std::setw of a double)."""
    assert checkpoint.format_block_index(*index) == text
    assert checkpoint.read_block_index(text, len(index) - 1) == index


def test_read_block_index_rank_check():
    with pytest.raises(ValueError):
        checkpoint.read_block_index("3:1-2", 1)


def test_lookup3_known_answers():
    """The metadata checksum is Bob Jenkins' lookup3 hashlittle; the values
    the published lookup3.c driver prints for its test strings."""
    assert hdf5.lookup3(b"") == 0xDEADBEEF
    assert hdf5.lookup3(b"", 0xDEADBEEF) == 0xBD5B7DDE
    assert hdf5.lookup3(b"Four score and seven years ago") == 0x17770551
    assert hdf5.lookup3(b"Four score and seven years ago", 1) == 0xCD628161


def _sample(tmp_path, rank=2):
    rng = np.random.default_rng(0)
    if rank == 2:
        blocks = [(2, i, j) for i in range(4) for j in range(4)]
        u = rng.normal(size=(16, 5, 8, 8))
    else:
        blocks = [(3, i) for i in range(8)]
        u = rng.normal(size=(8, 5, 16))
    path = str(tmp_path / "chkpt.0003.h5")
    checkpoint.write(path, blocks, u, iteration=12, time=0.25,
                     run_config={"cfl": 0.4, "depth": 2, "outdir": "data", "plm_theta": 1.5},
                     schedule={"checkpoint": {"next_due": 0.3, "count": 3}}, module="euler2d", git_commit="0123abc")
    return path, blocks, u


@pytest.mark.parametrize("rank", [1, 2])
def test_write_read_round_trip(tmp_path, rank):
    path, blocks, u = _sample(tmp_path, rank)
    with open(path, "rb") as f:
        assert f.read(8) == b"\x89HDF\r\n\x1a\n"
    ck = checkpoint.read(path, rank)
    assert ck["iteration"] == 12 and ck["time"] == 0.25 and ck["module"] == "euler2d"
    assert ck["git_commit"] == "0123abc"
    assert ck["run_config"] == {"cfl": 0.4, "depth": 2, "outdir": "data", "plm_theta": 1.5}
    assert ck["schedule"]["checkpoint"] == {"next_due": 0.3, "count": 3}
    assert ck["schedule"]["time_series"] == {"next_due": 0.0, "count": 0}
    for b, a in zip(blocks, u):
        assert np.array_equal(ck["conserved"][b], a)
    # the stored blocks use the reference's logical layout: [n(, n)] of n_c-double arrays
    raw = hdf5.read_file(path)["solution"]["conserved"]
    name = checkpoint.format_block_index(*blocks[5])
    assert raw[name].shape == u.shape[2:] + (5,)
    assert np.array_equal(raw[name][..., 2], u[5, 2])


def test_reader_rejects_a_corrupted_header(tmp_path):
    path, _, _ = _sample(tmp_path)
    b = bytearray(open(path, "rb").read())
    b[20] ^= 0x01                    # inside the superblock: its checksum fails
    open(path, "wb").write(bytes(b))
    with pytest.raises(hdf5.FormatError):
        hdf5.read_file(path)


H5DUMP = "/opt/conda/bin/h5dump"


@pytest.mark.skipif(not os.path.exists(H5DUMP), reason="no HDF5 library tools in this image")
def test_checkpoint_read_by_the_hdf5_library(tmp_path):
    """The HDF5 library itself (its h5dump tool) reads the checkpoint: the
    types the reference writes (H5T_ARRAY of n_c doubles per zone, the
    rational compound, C strings, native int and double), and every block's
    data exported by the library equals what was written, byte for byte."""
    path, blocks, u = _sample(tmp_path)
    head = subprocess.run([H5DUMP, "-H", path], capture_output=True, text=True, timeout=60)
    assert head.returncode == 0, head.stderr
    text = head.stdout
    assert 'GROUP "solution"' in text and 'GROUP "conserved"' in text and 'GROUP "run_config"' in text
    assert "H5T_ARRAY { [5] H5T_IEEE_F64LE }" in text
    assert 'H5T_STD_I64LE "num"' in text and 'H5T_STD_U64LE "den"' in text
    assert "H5T_STD_I32LE" in text and "CTYPE H5T_C_S1" in text
    for k in (0, 7, 15):
        name = checkpoint.format_block_index(*blocks[k])
        out = str(tmp_path / "block.bin")
        res = subprocess.run([H5DUMP, "-d", "/solution/conserved/" + name, "-b", "NATIVE", "-o", out, path],
                             capture_output=True, text=True, timeout=60)
        assert res.returncode == 0, res.stderr
        got = np.fromfile(out, dtype="<f8").reshape(8, 8, 5)
        assert np.array_equal(got, np.moveaxis(u[k], 0, -1))
    out = str(tmp_path / "it.bin")
    assert subprocess.run([H5DUMP, "-d", "/solution/iteration", "-b", "NATIVE", "-o", out, path],
                          capture_output=True, timeout=60).returncode == 0
    it = np.fromfile(out, dtype=[("num", "<i8"), ("den", "<u8")])
    assert it["num"][0] == 12 and it["den"][0] == 1


torch = pytest.importorskip("torch")


@pytest.mark.gpu
@pytest.mark.skipif(not torch.cuda.is_available(), reason="needs a HIP device")
def test_restart_continues_bitwise(tmp_path):
    """Checkpoint after 3 steps, restore into a fresh solver, 3 more steps:
    identical to 6 uninterrupted steps."""
    from mara3p_amd import problems
    from mara3p_amd.solvers import Euler2DSolver
    _, u0 = problems.kelvin_helmholtz(depth=1, block_size=64, seed=3)
    a = Euler2DSolver(1, 64)
    a.set_state(u0)
    dt = 0.4 * a.dl / 2.0
    for _ in range(3):
        a.step(dt)
    path = str(tmp_path / "chkpt.0000.h5")
    checkpoint.save_solver(path, a, iteration=3, time=3 * dt, module="euler2d")
    for _ in range(3):
        a.step(dt)
    b = Euler2DSolver(1, 64)
    it, t = checkpoint.restore_solver(path, b)
    assert it == 3 and t == 3 * dt
    for _ in range(3):
        b.step(dt)
    assert np.array_equal(a.state(), b.state())
