"""GPU checks of solver contracts beyond plain parity, through the C ABI:

* cg_solver::solve(A, x, b, precond) takes A apart from the preconditioner
  (krylov/cg.cpp:136-223): lal_pcg_solve_op_f64 runs the Krylov products on
  the caller's matrix and the V-cycle of a hierarchy built for another one;
* the FP32-V-cycle scheme scales the residual it hands to the FP32 levels, so
  right-hand sides far outside FP32's range solve like any other;
* the differentiable solve refuses an adjoint solve that did not converge.
"""
import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import op_precond_matrix
from tests.gpu_util import bits_equal, dev, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def _dev_csr(h, rp, ci, v, m, descr=True):
    RP, CI, V = dev(rp, np.int32), dev(ci, np.int32), dev(v, np.float64)
    d = lal.CsrDescr(h, m, m, len(ci), RP, CI) if descr else None
    return RP, CI, V, (RP, CI, V, m, len(ci), d)


def test_pcg_op_matches_reference_golden(h, golden):
    """Hierarchy for 8^3 log-normal diffusion, Krylov products on 8^3
    Poisson: the reference's iteration count and solution (cg_solver::solve
    with the composed SA-AMG preconditioner, golden_cg.npz)."""
    g = golden("cg")
    rp, ci, v, m = pb.poisson3d(8)
    drp, dci, dv, _ = op_precond_matrix()
    DRP, DCI, DV = dev(drp, np.int32), dev(dci, np.int32), dev(dv, np.float64)
    H = lal.Hierarchy(h, m, len(dci), DRP, DCI, DV)
    for descr in (True, False):
        log(f"pcg_op golden 8^3 descr={descr}")
        _, _, _, A = _dev_csr(h, rp, ci, v, m, descr)
        B = dev(g["op_b"])
        X = torch.zeros(m, dtype=torch.float64, device="cuda")
        info = H.pcg_op(A, B, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
        x = host(X)
        it_ref = int(g["op_it"][0])
        assert info.converged and abs(info.iterations - it_ref) <= 1, (info.iterations, it_ref)
        err = np.linalg.norm(x - g["op_x"]) / np.linalg.norm(g["op_x"])
        assert err <= 1e-10, err


def test_pcg_op_against_oracle_32(h, oracle):
    """32^3: hierarchy for a mildly log-normal diffusion operator (the
    situation of a hierarchy reused while the values drift), Krylov products
    on Poisson; the oracle's iteration count within one and its solution
    within 1e-10 relative (the north-star bar for the FP64 solve)."""
    nx = 32
    rp, ci, v, m = pb.poisson3d(nx)
    k = np.exp(0.25 * pb.normal(m, seed=21))
    drp, dci, dv, _ = pb.diffusion3d(nx, cell_kappa=k)
    DRP, DCI, DV = dev(drp, np.int32), dev(dci, np.int32), dev(dv, np.float64)
    H = lal.Hierarchy(h, m, len(dci), DRP, DCI, DV)
    ho = oracle.saamg_setup(CSR(drp, dci, dv, m, m))
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=22))
    xo, ito, _ = ho.pcg_op(CSR(rp, ci, v, m, m), b, rel_tol=1e-10)
    _, _, _, A = _dev_csr(h, rp, ci, v, m)
    X = torch.zeros(m, dtype=torch.float64, device="cuda")
    log("pcg_op 32^3 vs oracle")
    info = H.pcg_op(A, dev(b), X, rel_tol=1e-10, abs_tol=1e-300, max_iter=1000)
    x = host(X)
    r = b - pb.rhs_from_solution(rp, ci, v, x)
    err = np.linalg.norm(x - xo) / np.linalg.norm(xo)
    res = np.abs(r).max() / np.abs(b).max()
    log(f"pcg_op 32^3: {info.iterations} iterations (oracle {ito}), rel. error {err:.2e}, "
        f"rel. residual {res:.2e}")
    assert info.converged and abs(info.iterations - ito) <= 1, (info.iterations, ito)
    assert err <= 1e-10
    assert res <= 1.01e-10


def test_pcg_op_scaled_hierarchy_gives_own_iterates(h):
    """A hierarchy built for 4A: strength, aggregation and the prolongator's
    pattern and values do not change when A is multiplied by 4, every coarse
    operator (and the coarsest Cholesky factor, by 2) scales exactly, so its
    V-cycle is exactly a quarter of A's.  PCG is invariant to that, so
    lal_pcg_solve_op_f64(A, hier(4A)) reproduces lal_pcg_solve_f64(hier(A))
    bit for bit.  A solve that used the hierarchy's own matrix (4A) would
    return x/4."""
    rp, ci, v, m = pb.poisson3d(24)
    RP, CI, V, A = _dev_csr(h, rp, ci, v, m)
    H1 = lal.Hierarchy(h, m, len(ci), RP, CI, V)
    H2 = lal.Hierarchy(h, m, len(ci), RP, CI, V * 4.0)
    b = dev(pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=3)))
    X1 = torch.zeros(m, dtype=torch.float64, device="cuda")
    X2 = torch.zeros_like(X1)
    log("pcg_op 4A vs A")
    i1 = H1.pcg(b, X1, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    i2 = H2.pcg_op(A, b, X2, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    assert i1.converged and i2.converged and i1.iterations == i2.iterations
    assert bits_equal(host(X1), host(X2))


def test_pcg_op_refuses_size_mismatch(h):
    rp, ci, v, m = pb.poisson3d(8)
    RP, CI, V, _ = _dev_csr(h, rp, ci, v, m)
    H = lal.Hierarchy(h, m, len(ci), RP, CI, V)
    rp2, ci2, v2, m2 = pb.poisson3d(9)
    _, _, _, A2 = _dev_csr(h, rp2, ci2, v2, m2)
    X = torch.zeros(m2, dtype=torch.float64, device="cuda")
    with pytest.raises(lal.LalError):
        H.pcg_op(A2, X.clone(), X)


def test_fp32_vcycle_scheme_far_outside_fp32_range(h):
    """b scaled by 2^-140 (below FP32's smallest subnormal) and by 2^+120
    (past FP32's largest value when squared partials are formed): the FP32
    V-cycle receives r scaled by a power of two chosen from ||r0||, so every
    FP64 iterate is the unscaled solve's times the same power of two, bit
    for bit, and the iteration count is the same."""
    rp, ci, v, m = pb.poisson3d(20)
    RP, CI, V, _ = _dev_csr(h, rp, ci, v, m)
    H = lal.Hierarchy(h, m, len(ci), RP, CI, V, build_f32=1)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=4))
    X = torch.zeros(m, dtype=torch.float64, device="cuda")
    log("pcg_mixed unscaled")
    i0 = H.pcg_mixed(dev(b), X, rel_tol=1e-10, abs_tol=1e-300, max_iter=200)
    x0 = host(X)
    assert i0.converged
    for e in (-140, 120):
        s = np.ldexp(1.0, e)
        X.zero_()
        log(f"pcg_mixed b * 2^{e}")
        i1 = H.pcg_mixed(dev(b * s), X, rel_tol=1e-10, abs_tol=1e-300, max_iter=200)
        assert i1.converged and i1.iterations == i0.iterations, (e, i1.iterations, i0.iterations)
        assert bits_equal(host(X), x0 * s), e


def test_saamg_solve_refuses_unconverged_adjoint():
    """The adjoint solve of a loss whose gradient holds a NaN cannot
    converge: backward raises instead of returning a gradient."""
    from linalglib_amd.torch_solve import saamg_solve

    rp, ci, v, m = pb.poisson3d(8)
    RP, CI = dev(rp, np.int32), dev(ci, np.int32)
    val = dev(v).requires_grad_(True)
    b = dev(pb.uniform(m, seed=5))
    x = saamg_solve(val, b, RP, CI, tol=1e-10, max_iter=60)
    w = torch.ones(m, dtype=torch.float64, device="cuda")
    w[3] = float("nan")
    loss = (x * w).sum()
    with pytest.raises(lal.LalError):
        loss.backward()


def test_saamg_solve_reuses_hierarchy(h, oracle):
    """saamg_solve with a hierarchy built for other values of the same
    pattern: the forward solution and dL/db, dL/dA agree with the oracle's
    FP64 solves to 1e-8 relative."""
    from linalglib_amd.torch_solve import saamg_solve

    nx = 12
    rp, ci, v, m = pb.poisson3d(nx)
    k = np.exp(0.5 * pb.normal(m, seed=6))
    drp, dci, dv, _ = pb.diffusion3d(nx, cell_kappa=k)
    assert np.array_equal(drp, rp) and np.array_equal(dci, ci)
    RP, CI = dev(rp, np.int32), dev(ci, np.int32)
    H = lal.Hierarchy(lal.Handle(), m, len(ci), RP, CI, dev(v))
    val = dev(dv).requires_grad_(True)
    b = dev(pb.uniform(m, seed=7)).requires_grad_(True)
    xt = dev(pb.normal(m, seed=8))
    x = saamg_solve(val, b, RP, CI, tol=1e-12, hierarchy=H)
    loss = 0.5 * ((x - xt) ** 2).sum()
    loss.backward()
    A = CSR(rp, ci, dv, m, m)
    ho = oracle.saamg_setup(A)
    xo, _, _ = ho.pcg(host(b.detach()), rel_tol=1e-13)
    lo, _, _ = ho.pcg(xo - host(xt), rel_tol=1e-13)
    rows = np.repeat(np.arange(m), np.diff(rp))
    dA = -lo[rows] * xo[ci]
    xg = host(x.detach())
    assert np.linalg.norm(xg - xo) / np.linalg.norm(xo) <= 1e-8
    assert np.linalg.norm(host(b.grad) - lo) / np.linalg.norm(lo) <= 1e-8
    assert np.linalg.norm(host(val.grad) - dA) / np.linalg.norm(dA) <= 1e-8


def test_pcg_graphed_preconditioner_replays_match(h, oracle):
    """On a small finest level (2D Poisson 128^2) the PCG's V-cycle is run
    eagerly on its first application, captured as a HIP graph on its second
    and replayed from then on, with the solver's own eager kernels and host
    reads between replays.  A second solve, every application a replay,
    reproduces the first bit for bit, in FP64 and in the reliable-update FP32
    scheme, and the FP64 solve takes the oracle's iterations (within one)."""
    rp, ci, v, m = pb.poisson2d(128)
    RP, CI, V, _ = _dev_csr(h, rp, ci, v, m)
    H = lal.Hierarchy(h, m, len(ci), RP, CI, V, build_f32=1)
    b_h = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=11))
    b = dev(b_h)
    for name in ("pcg", "pcg_reliable"):
        xs, its = [], []
        for _ in range(2):
            X = torch.zeros(m, dtype=torch.float64, device="cuda")
            info = getattr(H, name)(b, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
            torch.cuda.synchronize()
            assert info.converged, name
            xs.append(host(X))
            its.append(info.iterations)
        log(f"graphed preconditioner, {name}: {its[0]} / {its[1]} iterations")
        assert its[0] == its[1] and bits_equal(xs[0], xs[1]), name
        if name == "pcg":
            _, ito, _ = oracle.saamg_setup(CSR(rp, ci, v, m, m)).pcg(b_h, rel_tol=1e-10)
            assert abs(its[0] - ito) <= 1, (its[0], ito)
