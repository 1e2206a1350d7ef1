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
