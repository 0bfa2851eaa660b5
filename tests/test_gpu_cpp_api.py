"""Runs the host C++ test of include/lal_linalg.hpp (LinAlgLib's API names over
the C ABI) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_linalg_api")


def test_cpp_api_program():
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (python -m linalglib_amd.build)")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=100)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL OK" in r.stdout
