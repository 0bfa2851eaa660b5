"""LinAlgLib's own front end on the MI355X backend: runs the program of
tests/dropin (the reference's library sources with the device shim bodies of
tests/dropin/lal_device_shims.cpp, linked with liblal.so), whose checks
compare each device-resident call of the reference's public API -- SpMV,
residual, BLAS-1, scan, diagonal, SpGEMM, transpose, strength, the PMIS shims,
cg_solver with Jacobi and SA-AMG preconditioning, bicgstab_solver -- with the
same call on the reference's host path.  The program is built where the
reference sources are (make -C tests/dropin) and travels with the tree."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "oracle", "_ref", "dropin", "dropin_test")


def test_reference_front_end_on_the_device():
    if not os.path.exists(BIN):
        pytest.skip("tests/dropin not built (needs the reference sources: make -C tests/dropin)")
    # One OpenMP thread for the reference's host path: its parallel reductions
    # combine the threads' partial sums in the order the threads finish, so
    # with several threads the host BiCGSTAB's iteration count moves between
    # runs of the same binary (62 and 68 seen against the device's 67), and
    # the host-against-device comparison needs a fixed host result.
    env = dict(os.environ, OMP_NUM_THREADS="1")
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=200, env=env)
    print(r.stdout[-6000:])
    assert r.returncode == 0, r.stdout[-6000:] + r.stderr[-2000:]
    assert "DROPIN OK" in r.stdout
    assert "FAIL " not in r.stdout
