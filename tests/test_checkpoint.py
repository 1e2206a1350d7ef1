"""Checkpoints in the reference's layout (src/app_problem.cpp:55-78,
src/module_euler_io.cpp:81-102, block names by mesh::format_block_index,
src/mesh_block_index.hpp:353-408)."""
import numpy as np
import pytest

from mara3p_amd import checkpoint


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


def test_write_read_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    blocks = [(2, i, j) for i in range(4) for j in range(4)]
    u = rng.normal(size=(16, 5, 8, 8))
    path = str(tmp_path / "chkpt.0003.npz")
    checkpoint.write(path, blocks, u, iteration=12, time=0.25, run_config={"cfl": 0.4, "depth": 2},
                     schedule={"checkpoint_count": 3}, module="euler2d")
    ck = checkpoint.read(path, 2)
    assert ck["iteration"] == 12 and ck["time"] == 0.25 and ck["module"] == "euler2d"
    assert ck["run_config"]["cfl"] == 0.4 and ck["schedule"]["checkpoint_count"] == 3
    for b, a in zip(blocks, u):
        assert np.array_equal(ck["conserved"][b], a)
    # the stored datasets use the reference's logical layout [n, n, n_c]
    with np.load(path) as z:
        assert z["solution/conserved/2:1-3"].shape == (8, 8, 5)
        assert np.array_equal(z["solution/conserved/2:1-3"][2, 5], u[blocks.index((2, 1, 3)), :, 2, 5])


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
    path = str(tmp_path / "chkpt.0000.npz")
    checkpoint.save_solver(path, a, iteration=3, time=3 * dt, module="euler2d")
    for _ in range(3):
        a.step(dt)
    b = Euler2DSolver(1, 64)
    it, t = checkpoint.restore_solver(path, b)
    assert it == 3 and t == 3 * dt
    for _ in range(3):
        b.step(dt)
    assert np.array_equal(a.state(), b.state())
