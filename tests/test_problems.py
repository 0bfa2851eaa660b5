"""The synthetic problem generators (linalglib_amd/problems.py) against
independent constructions, and the Python binding's argument checks that run
before any device call.  Host only."""
import numpy as np
import pytest
import scipy.sparse as sp

from linalglib_amd import problems as pb


def _csr(rp, ci, v, m, n=None):
    return sp.csr_matrix((v, ci, rp), shape=(m, m if n is None else n))


def _sorted_rows(rp, ci):
    return all(np.all(np.diff(ci[rp[i]: rp[i + 1]]) > 0) for i in range(len(rp) - 1))


def test_poisson_matches_kronecker_sums():
    for nx in (1, 2, 5):
        T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(nx, nx))
        I = sp.eye(nx)
        rp, ci, v, m = pb.poisson2d(nx)
        assert abs(_csr(rp, ci, v, m) - (sp.kron(I, T) + sp.kron(T, I))).max() == 0
        assert _sorted_rows(rp, ci)
        rp, ci, v, m = pb.poisson3d(nx)
        L3 = sp.kron(I, sp.kron(I, T)) + sp.kron(I, sp.kron(T, I)) + sp.kron(T, sp.kron(I, I))
        assert abs(_csr(rp, ci, v, m) - L3).max() == 0
        assert _sorted_rows(rp, ci)


@pytest.mark.parametrize("wind", [(20.0, -10.0), (-3.0, 7.0), (0.0, 0.0)])
def test_convection_diffusion_upwind(wind):
    """-lap u + w . grad u with first-order upwinding: backward differences
    for a positive wind component, forward for a negative one."""
    n = 7
    h = 1.0 / (n + 1)
    rp, ci, v, m = pb.convection_diffusion2d(n, wind)
    T = sp.diags([-1, 2, -1], [-1, 0, 1], shape=(n, n)) / h**2
    Bk = sp.diags([-1, 1], [-1, 0], shape=(n, n)) / h
    Fw = sp.diags([-1, 1], [0, 1], shape=(n, n)) / h
    I = sp.eye(n)
    Dx = Bk if wind[0] > 0 else Fw
    Dy = Bk if wind[1] > 0 else Fw
    ref = sp.kron(I, T) + sp.kron(T, I) + wind[0] * sp.kron(I, Dx) + wind[1] * sp.kron(Dy, I)
    A = _csr(rp, ci, v, m)
    assert abs(A - ref).max() <= 1e-12 * abs(ref).max()
    assert _sorted_rows(rp, ci)
    # M-matrix pattern: positive diagonal, non-positive off-diagonals,
    # weak diagonal dominance in every row
    Ad = A.toarray()
    off = Ad - np.diag(np.diag(Ad))
    assert np.all(np.diag(Ad) > 0) and np.all(off <= 0)
    assert np.all(np.diag(Ad) + off.sum(axis=1) >= -1e-9 * np.diag(Ad))


def test_rhs_from_solution_is_a_times_x():
    rp, ci, v, m = pb.poisson3d(6)
    x = pb.uniform(m, seed=3)
    assert np.allclose(pb.rhs_from_solution(rp, ci, v, x), _csr(rp, ci, v, m) @ x,
                       rtol=0, atol=1e-13)


def test_krylov_binding_refuses_bad_preconditioners_before_any_device_call():
    from linalglib_amd import lal

    A = (None, None, None, 4, 4, None)
    with pytest.raises(ValueError):
        lal._krylov_args(A, "ilu", None)
    with pytest.raises(ValueError):
        lal._krylov_args(A, "saamg", None)
