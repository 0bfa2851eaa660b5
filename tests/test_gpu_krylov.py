"""BiCGSTAB and restarted GMRES on the device (lal_bicgstab_solve,
lal_gmres_solve) with no, Jacobi and SA-AMG preconditioning, and the Jacobi
preconditioner (lal_jacobi_solve_f64).

The device inner products are tree reductions, so the iterates differ from the
reference's sequential sums in the last bits.  Parity therefore is: on the
reference's golden fixtures (BiCGSTAB), its iteration count within 2 and its
solution within 1e-8 relative (the solves run to 1e-10); against the oracle
(nonsymmetric systems, GMRES), its iteration count within a few percent and
a solution as close to the direct solution, and a true residual as small, as
the oracle's own (within 10x)."""
import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import grid
from tests.gpu_util import DevCSR, bits_equal, dev, host, log

pytestmark = pytest.mark.gpu

PC = ["none", "jacobi", "saamg"]


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


class DevSystem:
    """A square matrix on the device with its SpMV descriptor (and, on
    request, its SA-AMG hierarchy)."""

    def __init__(self, h, A, hierarchy=False, exact=1):
        self.d = DevCSR(A)
        self.descr = lal.CsrDescr(h, A.m, A.m, A.nnz, self.d.row_ptr, self.d.col_ind)
        self.A = (self.d.row_ptr, self.d.col_ind, self.d.val, A.m, A.nnz, self.descr)
        self.H = (lal.Hierarchy(h, A.m, A.nnz, self.d.row_ptr, self.d.col_ind, self.d.val,
                                exact=exact) if hierarchy else None)


def convection_diffusion(n):
    rp, ci, v, m = pb.convection_diffusion2d(n)
    return CSR(rp, ci, v, m, m)


def rel_err(x, y):
    return np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)


def true_rel_res(A, x, b):
    import scipy.sparse as sp

    S = sp.csr_matrix((A.val, A.col_ind, A.row_ptr), shape=(A.m, A.n))
    return np.abs(b - S @ x).max() / np.abs(b).max()


def direct(A, b):
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla

    return spla.spsolve(sp.csc_matrix(sp.csr_matrix((A.val, A.col_ind, A.row_ptr),
                                                    shape=(A.m, A.n))), b)


def as_accurate_as_oracle(A, b, x, xo):
    """The device solution is as close to the direct solution as the oracle's
    (within 10x), and its true residual as small (within 10x): the two solves
    stop at the same tolerance but round differently."""
    xd = direct(A, b)
    assert rel_err(x, xd) <= 10 * rel_err(xo, xd) + 1e-12, (rel_err(x, xd), rel_err(xo, xd))
    assert true_rel_res(A, x, b) <= 10 * true_rel_res(A, xo, b) + 1e-14


def test_jacobi_solve_bitwise(h):
    rng = np.random.default_rng(0)
    for n in (1, 7, 1000, 300001):
        log(f"jacobi_solve n={n}")
        r = rng.standard_normal(n)
        d = rng.uniform(0.5, 4, n)
        x = torch.zeros(n, dtype=torch.float64, device="cuda")
        lal.jacobi_solve(h, dev(r), dev(d), x)
        assert bits_equal(host(x), r / d)


@pytest.mark.parametrize("name", ["2d8", "3d8"])
def test_bicgstab_matches_reference_golden(h, golden, name):
    g = golden("krylov")
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    S = DevSystem(h, A, hierarchy=True)
    b = g[f"{name}_b"]
    for pc in range(3):
        log(f"bicgstab {name} precond={PC[pc]}")
        x = torch.zeros(m, dtype=torch.float64, device="cuda")
        info = lal.bicgstab(h, S.A, dev(b), x, PC[pc], S.H, rel_tol=1e-10, abs_tol=1e-300,
                            max_iter=500)
        it_ref = int(g[f"{name}_p{pc}_it"][0])
        assert abs(info.iterations - it_ref) <= 2, (pc, info.iterations, it_ref)
        assert info.converged
        assert rel_err(host(x), g[f"{name}_p{pc}_x"]) <= 1e-8, pc


@pytest.mark.parametrize("pc", [0, 1, 2])
def test_bicgstab_nonsymmetric_matches_oracle(h, oracle, pc):
    A = convection_diffusion(48)
    S = DevSystem(h, A, hierarchy=(pc == 2))
    rng = np.random.default_rng(3 + pc)
    b = rng.uniform(-1, 1, A.m)
    x0 = rng.uniform(-0.1, 0.1, A.m)
    log(f"bicgstab convection-diffusion 48^2 precond={PC[pc]}")
    ho = oracle.saamg_setup(A) if pc == 2 else None
    xo, ko = oracle.bicgstab(A, b, x0=x0, precond=pc, hier=ho, rel_tol=1e-10, max_iter=2000)
    x = dev(x0)
    info = lal.bicgstab(h, S.A, dev(b), x, PC[pc], S.H, rel_tol=1e-10, abs_tol=1e-300,
                        max_iter=2000)
    assert info.converged and abs(info.iterations - ko) <= max(3, ko // 10), \
        (info.iterations, ko)
    as_accurate_as_oracle(A, b, host(x), xo)


def test_bicgstab_breakdown_branch(h, oracle):
    """A = 2I: after one step r is exactly zero, t.r = 0 takes the breakdown
    branch and x = x0 + alpha p = b / 2 with zero iterations, as in
    bicgstab.cpp; the oracle agrees."""
    m = 100
    A = CSR(np.arange(m + 1, dtype=np.int32), np.arange(m, dtype=np.int32), np.full(m, 2.0),
            m, m)
    b = np.random.default_rng(1).uniform(-1, 1, m)
    xo, ko = oracle.bicgstab(A, b, rel_tol=1e-12, max_iter=10)
    assert ko == 0 and bits_equal(xo, b / 2)
    S = DevSystem(h, A)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    info = lal.bicgstab(h, S.A, dev(b), x, "none", None, rel_tol=1e-12, max_iter=10)
    assert info.iterations == 0 and not info.converged
    assert bits_equal(host(x), b / 2)


@pytest.mark.parametrize("pc,restart", [(0, 8), (0, 30), (1, 20), (2, 10)])
def test_gmres_matches_oracle(h, oracle, pc, restart):
    A = convection_diffusion(40)
    S = DevSystem(h, A, hierarchy=(pc == 2))
    rng = np.random.default_rng(11 + pc)
    b = rng.uniform(-1, 1, A.m)
    log(f"gmres({restart}) convection-diffusion 40^2 precond={PC[pc]}")
    ho = oracle.saamg_setup(A) if pc == 2 else None
    xo, ko = oracle.gmres(A, b, restart=restart, precond=pc, hier=ho, rel_tol=1e-10,
                          max_iter=3000)
    x = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    info = lal.gmres(h, S.A, dev(b), x, restart, PC[pc], S.H, rel_tol=1e-10, abs_tol=1e-300,
                     max_iter=3000)
    assert info.converged and abs(info.iterations - ko) <= max(1, ko // 20), \
        (info.iterations, ko)
    as_accurate_as_oracle(A, b, host(x), xo)


def test_gmres_exact_in_one_cycle(h):
    """With restart >= m - 1 GMRES is a direct method: one cycle of m - 1
    Arnoldi steps plus the final correction reaches the direct solution."""
    import scipy.sparse as sp

    m = 40
    M = (sp.random(m, m, density=0.15, random_state=7, format="csr") + sp.eye(m) * 3).tocsr()
    M.sort_indices()
    A = CSR(M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data, m, m)
    b = np.random.default_rng(7).uniform(-1, 1, m)
    xd = np.linalg.solve(M.toarray(), b)
    S = DevSystem(h, A)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    lal.gmres(h, S.A, dev(b), x, m - 1, "none", None, rel_tol=1e-14, abs_tol=1e-300,
              max_iter=3)
    assert rel_err(host(x), xd) <= 1e-10


def test_krylov_saamg_poisson(h, oracle):
    """SA-AMG-preconditioned BiCGSTAB and GMRES on 3D Poisson 32^3 in the fast
    (non-exact) mode: the oracle's iteration counts within 2 and the true
    residual at the tolerance."""
    rp, ci, v, m = pb.poisson3d(32)
    A = CSR(rp, ci, v, m, m)
    S = DevSystem(h, A, hierarchy=True, exact=0)
    ho = oracle.saamg_setup(A)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=4))
    log("bicgstab saamg poisson 32^3")
    xo, ko = oracle.bicgstab(A, b, precond=2, hier=ho, rel_tol=1e-10, max_iter=500)
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    info = lal.bicgstab(h, S.A, dev(b), x, "saamg", S.H, rel_tol=1e-10, abs_tol=1e-300,
                        max_iter=500)
    assert info.converged and abs(info.iterations - ko) <= 2, (info.iterations, ko)
    assert true_rel_res(A, host(x), b) <= 10 * true_rel_res(A, xo, b) + 1e-14
    log("gmres(20) saamg poisson 32^3")
    xo, ko = oracle.gmres(A, b, restart=20, precond=2, hier=ho, rel_tol=1e-10, max_iter=100)
    x.zero_()
    info = lal.gmres(h, S.A, dev(b), x, 20, "saamg", S.H, rel_tol=1e-10, abs_tol=1e-300,
                     max_iter=100)
    assert info.converged and abs(info.iterations - ko) <= 1, (info.iterations, ko)
    assert true_rel_res(A, host(x), b) <= 10 * true_rel_res(A, xo, b) + 1e-14


def test_krylov_refuses_bad_arguments(h):
    rp, ci, v, m = pb.poisson2d(4)
    S = DevSystem(h, CSR(rp, ci, v, m, m))
    b = torch.ones(m, dtype=torch.float64, device="cuda")
    x = torch.zeros(m, dtype=torch.float64, device="cuda")
    rp_, ci_, val, m_, nnz, descr = S.A
    with pytest.raises(lal.LalError):  # SA-AMG preconditioning without a hierarchy
        lal.call("lal_bicgstab_solve", h.h, descr.d, m, nnz, lal.ptr(rp_), lal.ptr(ci_),
                 lal.ptr(val), lal.ptr(b), lal.ptr(x), 2, None, None, None)
    with pytest.raises(lal.LalError):  # unknown preconditioner kind
        lal.call("lal_gmres_solve", h.h, descr.d, m, nnz, lal.ptr(rp_), lal.ptr(ci_),
                 lal.ptr(val), lal.ptr(b), lal.ptr(x), 10, 3, None, None, None)
    with pytest.raises(lal.LalError):  # restart < 1
        lal.call("lal_gmres_solve", h.h, descr.d, m, nnz, lal.ptr(rp_), lal.ptr(ci_),
                 lal.ptr(val), lal.ptr(b), lal.ptr(x), 0, 0, None, None, None)
    with pytest.raises(ValueError):
        lal.bicgstab(h, S.A, b, x, "saamg", None)
    other = DevSystem(h, CSR(*pb.poisson2d(5)[:3], 25, 25), hierarchy=True)
    with pytest.raises(lal.LalError):  # hierarchy built for another matrix
        lal.bicgstab(h, S.A, b, x, "saamg", other.H)
