"""The C++ host API (reference-style module entry points) against the oracle,
run as the compiled program tests/cpp/test_host_api."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_host_api")


def test_host_api_program_is_built():
    assert os.path.exists(BIN), "run mara3p_amd/build.py"


@pytest.mark.gpu
def test_host_api_against_oracle():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    res = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(res.stdout)
    assert res.returncode == 0 and "ALL PASS" in res.stdout


@pytest.mark.gpu
def test_host_api_updated_solution_quadtree_against_amr_oracle(tmp_path):
    """euler2d::updated_solution of the C++ host API on a 2:1-balanced
    quad-tree with level jumps both ways (a level-1 block beside level-2
    blocks, four level-3 blocks): the program writes the tree, the input and
    the output; the output must equal the quad-tree oracle's stage
    (amr::extend_block over amr::get_or_create_block) bit for bit."""
    import numpy as np
    from oracle import amr as oamr
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("needs a HIP device")
    path = str(tmp_path / "amr.bin")
    res = subprocess.run([BIN, "--dump-amr", path], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout
    raw = open(path, "rb").read()
    nleaf, n = np.frombuffer(raw[:16], dtype=np.int64)
    idx = np.frombuffer(raw[16:16 + 24 * nleaf], dtype=np.int64).reshape(nleaf, 3)
    data = np.frombuffer(raw[16 + 24 * nleaf:], dtype=np.float64).reshape(2, nleaf, n, n, 5)
    leaves = [tuple(int(v) for v in row) for row in idx]
    assert leaves == sorted(leaves)
    assert {b[0] for b in leaves} == {1, 2, 3}
    levels = np.array([b[0] for b in leaves])
    ref = oamr.stage(np.ascontiguousarray(data[0]), leaves, levels, int(n), 1e-3, 1.5, 1.4, 0)
    assert np.array_equal(data[1], ref)
