"""Runs the host C++ test of include/lal_linalg.hpp (LinAlgLib's API names over
the C ABI) on the GPU, and checks the solution of its non-preconditioned CG
(cg_solver::solve_nonprecond, cg.cpp:60-134) bit for bit against the
reference's (tests/golden/golden_cg.npz, written by LinAlgLib's compiled
cg_solver on the same problem)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "test_linalg_api")


def test_cpp_api_program(golden):
    if not os.path.exists(BIN):
        pytest.fail(f"{BIN} not built (python -m linalglib_amd.build)")
    with tempfile.TemporaryDirectory() as tmp:
        out = os.path.join(tmp, "cg_x.bin")
        r = subprocess.run([BIN, out], capture_output=True, text=True, timeout=100)
        print(r.stdout)
        assert r.returncode == 0, r.stdout + r.stderr
        assert "ALL OK" in r.stdout
        x = np.fromfile(out, dtype=np.float64)
    g = golden("cg")
    assert "non-preconditioned CG %d iterations" % int(g["ones6_it"][0]) in r.stdout
    assert x.shape == g["ones6_x"].shape
    assert np.array_equal(x.view(np.int64), g["ones6_x"].view(np.int64))
