"""Unsmoothed-aggregation hierarchies built on the device (lal_uaamg_setup)
against LinAlgLib's uaamg_setup outputs (golden fixtures) bit for bit, and
solves on them against the CPU oracle."""
import hashlib

import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import grid
from tests.gpu_util import DevCSR, bits_equal, dev, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def _hier(h, A, **kw):
    d = DevCSR(A)
    H = lal.Hierarchy(h, A.m, A.nnz, d.row_ptr, d.col_ind, d.val, unsmoothed=True, **kw)
    H._keep = d
    return H


@pytest.mark.parametrize("name,full", [("2d8", True), ("3d8", True), ("2d128", False),
                                       ("3d16", False)])
def test_uaamg_device_matches_reference_golden(h, golden, name, full):
    g = golden("ua")
    rp, ci, v, m = grid(name)
    log(f"uaamg {name}")
    H = _hier(h, CSR(rp, ci, v, m, m))
    assert H.levels == int(g[f"{name}_levels"][0])
    for lvl in range(H.levels + 1):
        for kind in ("A", "P", "R") if lvl < H.levels else ("A",):
            grp, gci, gv, mm, nn = H.level(lvl, kind)
            key = f"{name}_{kind}{lvl}"
            assert np.array_equal([mm, nn, len(gci)], g[key + "_dims"]), key
            if full:
                assert np.array_equal(grp, g[key + "_rp"]) and np.array_equal(gci, g[key + "_ci"])
                assert bits_equal(gv, g[key + "_v"]), key
            else:
                s = (hashlib.sha256(np.ascontiguousarray(grp, np.int32).tobytes()).hexdigest()
                     + hashlib.sha256(np.ascontiguousarray(gci, np.int32).tobytes()).hexdigest()
                     + hashlib.sha256(np.ascontiguousarray(gv, np.float64).tobytes()).hexdigest())
                assert s.encode() == g[key + "_sha"].tobytes(), key


def test_uaamg_solves_match_oracle(h, oracle):
    """PCG with the unsmoothed-aggregation V-cycle, and amg_solve with
    Gauss-Seidel cycles in exact mode, against the oracle on the same
    hierarchy."""
    rp, ci, v, m = pb.poisson3d(20)
    A = CSR(rp, ci, v, m, m)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=3))
    ho = oracle.uaamg_setup(A)
    log("uaamg pcg 20^3")
    H = _hier(h, A)
    B = dev(b)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    info = H.pcg(B, x, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    xo, ito, _ = ho.pcg(b, rel_tol=1e-10, coarse_solver=1)
    assert info.converged and abs(info.iterations - ito) <= 1, (info.iterations, ito)
    assert np.linalg.norm(host(x) - xo) <= 1e-9 * np.linalg.norm(xo)
    log("uaamg amg_solve exact 20^3")
    He = _hier(h, A, exact=1)
    x.zero_()
    info = He.amg_solve(B, x, 1, 1, "v", "gs", rel_tol=1e-8, abs_tol=1e-300, max_iter=40)
    xo, ko, _ = ho.amg_solve(b, cycle=0, smoother=1, rel_tol=1e-8, max_cycle=40)
    assert info.iterations == ko and bits_equal(host(x), xo)
