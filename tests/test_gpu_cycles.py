"""AMG as a stand-alone solver on the device (lal_amg_solve): V, W and F
cycles with the Jacobi, Gauss-Seidel, symmetric Gauss-Seidel, SOR and SSOR
smoothers.  In exact mode the device solve is bit-identical to the
reference's (golden fixtures) and to the CPU oracle; in the fast mode it
converges like the oracle."""
import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import grid
from tests.gpu_util import DevCSR, bits_equal, dev, host, log

pytestmark = pytest.mark.gpu

CYC = ["v", "w", "f"]
SM = ["jacobi", "gs", "sgs", "sor", "ssor"]


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def _hier(h, A, **kw):
    d = DevCSR(A)
    H = lal.Hierarchy(h, A.m, A.nnz, d.row_ptr, d.col_ind, d.val, **kw)
    H._keep = d
    return H


@pytest.mark.parametrize("name", ["2d8", "3d8"])
def test_amg_solve_bitwise_reference_golden(h, golden, name):
    g = golden("cycles")
    rp, ci, v, m = grid(name)
    H = _hier(h, CSR(rp, ci, v, m, m), exact=1)
    B = dev(g[f"{name}_b"])
    for cyc in range(3):
        for sm in range(5):
            log(f"amg_solve {name} cycle={CYC[cyc]} smoother={SM[sm]}")
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            info = H.amg_solve(B, x, 1, 1, CYC[cyc], SM[sm], rel_tol=1e-9, abs_tol=1e-300,
                               max_iter=40)
            assert info.iterations == int(g[f"{name}_c{cyc}_s{sm}_k"][0]), (cyc, sm)
            assert bits_equal(host(x), g[f"{name}_c{cyc}_s{sm}_x"]), (cyc, sm)


def test_amg_solve_bitwise_oracle_sweeps(h, oracle):
    """n1 = 2, n2 = 1 on a diffusion operator with a nonzero initial guess."""
    rp, ci, v, m = grid("diff12")
    A = CSR(rp, ci, v, m, m)
    H = _hier(h, A, exact=1)
    ho = oracle.saamg_setup(A)
    rng = np.random.default_rng(4)
    b = rng.uniform(-1, 1, m)
    x0 = rng.uniform(-1, 1, m)
    B = dev(b)
    for cyc in range(3):
        for sm in range(5):
            log(f"amg_solve diff12 n1=2 cycle={CYC[cyc]} smoother={SM[sm]}")
            x = dev(x0)
            info = H.amg_solve(B, x, 2, 1, CYC[cyc], SM[sm], rel_tol=1e-8, abs_tol=1e-300,
                               max_iter=25)
            xo, ko, _ = ho.amg_solve(b, x0, n1=2, n2=1, cycle=cyc, smoother=sm, rel_tol=1e-8,
                                     max_cycle=25)
            assert info.iterations == ko, (cyc, sm, info.iterations, ko)
            assert bits_equal(host(x), xo), (cyc, sm)


def test_amg_solve_fast_mode_converges(h, oracle):
    """Fast kernels (sums reordered inside rows) on 3D Poisson 24^3: the same
    convergence as the oracle within a cycle or two."""
    rp, ci, v, m = pb.poisson3d(24)
    A = CSR(rp, ci, v, m, m)
    H = _hier(h, A)
    ho = oracle.saamg_setup(A)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=2))
    B = dev(b)
    for cyc in range(3):
        for sm in (0, 1, 4):
            log(f"amg_solve fast 24^3 cycle={CYC[cyc]} smoother={SM[sm]}")
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            info = H.amg_solve(B, x, 1, 1, CYC[cyc], SM[sm], rel_tol=1e-8, abs_tol=1e-300,
                               max_iter=60)
            xo, ko, _ = ho.amg_solve(b, n1=1, n2=1, cycle=cyc, smoother=sm, rel_tol=1e-8,
                                     max_cycle=60)
            assert info.converged == (ko <= 60)
            assert abs(info.iterations - ko) <= 2, (cyc, sm, info.iterations, ko)
            if info.converged:
                r = oracle.residual(A, host(x), b)
                assert np.abs(r).max() <= 1e-8 * np.abs(b).max() * 1.01


def test_amg_solve_argument_checks(h):
    rp, ci, v, m = grid("2d8")
    H = _hier(h, CSR(rp, ci, v, m, m))
    B = dev(np.ones(m))
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    for cyc, sm in ((3, 0), (0, 7), (-1, 0), (0, -1)):
        with pytest.raises(lal.LalError):
            lal.call("lal_amg_solve", h.h, H.hp, lal.ptr(B), lal.ptr(x), 1, 1, cyc, sm, None,
                     None)


def test_amg_solve_gauss_seidel_bitwise_oracle_large(h, oracle):
    """3D Poisson 40^3 (64000 rows, sets of ~540 rows on the finest level, so
    every level is swept by one workgroup and the finest level's x is larger
    than a workgroup's LDS): V and W cycles with Gauss-Seidel and SSOR in
    exact mode, bit-identical to the oracle after three cycles."""
    rp, ci, v, m = pb.poisson3d(40)
    A = CSR(rp, ci, v, m, m)
    H = _hier(h, A, exact=1)
    ho = oracle.saamg_setup(A)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=6))
    B = dev(b)
    for cyc in (0, 1):
        for sm in (1, 4):
            log(f"amg_solve 40^3 exact cycle={CYC[cyc]} smoother={SM[sm]}")
            x = torch.zeros(m, dtype=torch.float64, device="cuda")
            info = H.amg_solve(B, x, 1, 1, CYC[cyc], SM[sm], rel_tol=1e-300, abs_tol=1e-300,
                               max_iter=2)
            xo, ko, _ = ho.amg_solve(b, n1=1, n2=1, cycle=cyc, smoother=sm, rel_tol=1e-300,
                                     max_cycle=2)
            assert info.iterations == ko == 3, (info.iterations, ko)
            assert bits_equal(host(x), xo), (cyc, sm)
