"""Multicolour Gauss-Seidel smoothers (smoother 5: forward colour sweep, 6:
forward then backward) in the AMG cycles and as the PCG V-cycle's smoother.

These smoothers have no counterpart in LinAlgLib (whose Gauss-Seidel,
gauss_seidel.cpp:50-87, is the sequential sweep reproduced bit for bit by
smoother 1), so they are pinned by mathematics instead:

* on the 2D 5-point operator, greedy colouring in row order is the red-black
  colouring, and one V(1,0) cycle equals a red-black Gauss-Seidel sweep
  (written here from its textbook definition) followed by the coarse
  correction computed with numpy from the device hierarchy's own levels;
* the symmetric variant makes the V-cycle a symmetric positive definite
  operator, the forward-only one does not;
* AMG cycles converge at a rate close to the sequential sweep's, PCG with the
  symmetric multicolour V-cycle converges on the true residual, and one such
  V-cycle takes at most 2 ms on 3D Poisson 128^3.
"""
import time

import numpy as np
import pytest
import scipy.sparse as sp
import torch

from linalglib_amd import lal, problems as pb
from tests.gpu_util import bits_equal, dev, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def _hier(h, rp, ci, v, m, **opts):
    RP, CI, V = dev(rp, np.int32), dev(ci, np.int32), dev(v, np.float64)
    return lal.Hierarchy(h, m, len(ci), RP, CI, V, **opts), (RP, CI, V)


def _sparse(level):
    rp, ci, v, m, n = level
    return sp.csr_matrix((v, ci, rp), shape=(m, n))


def one_cycle(H, b, smoother, n1=1, n2=1, cycle="v"):
    """z = one cycle from x = 0 (amg_solve with max_iter = 0)."""
    X = torch.zeros(len(b), dtype=torch.float64, device="cuda")
    H.amg_solve(dev(b), X, n1, n2, cycle=cycle, smoother=smoother, rel_tol=1e-300,
                abs_tol=1e-300, max_iter=0)
    return host(X)


def test_red_black_sweep_closed_form(h):
    """2D Poisson 24^2 with one coarse level: V(1,0) with smoother 5 from
    x = 0 is the red-black sweep (red = (i + j) even first, each red row
    solved from its black neighbours, then each black row from the new reds),
    then x += P A_c^-1 R (b - A x) with the coarsest level solved by the
    reference's 1000 Gauss-Seidel sweeps."""
    nx = 24
    rp, ci, v, m = pb.poisson2d(nx)
    H, _ = _hier(h, rp, ci, v, m, max_levels=1)
    assert H.levels == 1
    b = pb.uniform(m, seed=31)
    z = one_cycle(H, b, "mcgs", n1=1, n2=0)
    A = _sparse(H.level(0, "A"))
    d = A.diagonal()
    off = A - sp.diags(d)
    f = np.arange(m)  # row f is grid point (f % nx, f // nx)
    red = (f % nx + f // nx) % 2 == 0
    x = np.zeros(m)
    for mask in (red, ~red):
        xn = (b - off @ x) / d
        x[mask] = xn[mask]
    R, P, Ac = (_sparse(H.level(0, "R")), _sparse(H.level(0, "P")),
                _sparse(H.level(1, "A")))
    bc = R @ (b - A @ x)
    ec = np.zeros(Ac.shape[0])
    Acd = Ac.toarray()
    for _ in range(1000):
        for j in range(len(ec)):
            ec[j] = (bc[j] - Acd[j] @ ec + Acd[j, j] * ec[j]) / Acd[j, j]
    x = x + P @ ec
    assert np.abs(z - x).max() <= 1e-12 * np.abs(x).max()


@pytest.mark.parametrize("gen", [lambda: pb.poisson3d(16),
                                 lambda: pb.diffusion3d(12, cell_kappa=np.exp(
                                     1.5 * pb.normal(12 ** 3, seed=7)))])
def test_symmetric_variant_is_spd_preconditioner(h, gen):
    rp, ci, v, m = gen()
    H, _ = _hier(h, rp, ci, v, m)
    r1, r2 = pb.uniform(m, seed=1), pb.uniform(m, seed=2)
    for smoother, symmetric in (("mcsgs", True), ("mcgs", False)):
        z1, z2 = one_cycle(H, r1, smoother), one_cycle(H, r2, smoother)
        a, b = z1 @ r2, r1 @ z2
        gap = abs(a - b) / abs(a)
        log(f"{smoother}: <M r1, r2> vs <r1, M r2> relative gap {gap:.2e}")
        if symmetric:
            assert gap <= 1e-12
            for s in range(3, 8):
                r = pb.uniform(m, seed=s)
                assert one_cycle(H, r, smoother) @ r > 0
        else:
            assert gap > 1e-8  # the forward-only V-cycle is not symmetric


@pytest.mark.parametrize("gen,name", [(lambda: pb.poisson3d(32), "poisson 32^3"),
                                      (lambda: pb.diffusion3d(24, cell_kappa=np.exp(
                                          1.0 * pb.normal(24 ** 3, seed=9))), "diffusion 24^3")])
def test_multicolour_cycles_converge_like_sequential(h, gen, name):
    """AMG as a solver, V(1,1): the multicolour sweeps take at most 1.5x the
    cycles of the sequential Gauss-Seidel sweep (and fewer than Jacobi);
    the true residual meets the tolerance."""
    rp, ci, v, m = gen()
    H, _ = _hier(h, rp, ci, v, m)
    b = pb.uniform(m, seed=5)
    counts = {}
    for sm in ("jacobi", "gs", "mcgs", "mcsgs"):
        X = torch.zeros(m, dtype=torch.float64, device="cuda")
        info = H.amg_solve(dev(b), X, 1, 1, cycle="v", smoother=sm, rel_tol=1e-8,
                           abs_tol=1e-300, max_iter=300)
        x = host(X)
        r = b - pb.rhs_from_solution(rp, ci, v, x)
        assert info.converged and np.abs(r).max() <= 1.01e-8 * np.abs(b).max(), sm
        counts[sm] = info.iterations
    log(f"{name}: cycles {counts}")
    assert counts["mcgs"] <= 1.5 * counts["gs"] + 1
    assert counts["mcgs"] < counts["jacobi"] and counts["mcsgs"] <= counts["mcgs"]


def test_pcg_with_symmetric_multicolour_vcycle(h, oracle):
    """FP64 PCG whose V-cycle smooths with symmetric multicolour Gauss-Seidel:
    true residual at the tolerance, solution within 1e-8 of the oracle's,
    in no more iterations than the Jacobi V-cycle's PCG."""
    from oracle.oracle import CSR

    rp, ci, v, m = pb.diffusion3d(32, cell_kappa=np.exp(1.0 * pb.normal(32 ** 3, seed=11)))
    H, _ = _hier(h, rp, ci, v, m)
    b = pb.normal(m, seed=12)
    X = torch.zeros(m, dtype=torch.float64, device="cuda")
    ij = H.pcg(dev(b), X, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    X.zero_()
    im = H.pcg(dev(b), X, rel_tol=1e-10, abs_tol=1e-300, max_iter=500, smoother="mcsgs")
    x = host(X)
    r = b - pb.rhs_from_solution(rp, ci, v, x)
    assert im.converged and np.abs(r).max() <= 1.01e-10 * np.abs(b).max()
    xo, _, _ = oracle.saamg_setup(CSR(rp, ci, v, m, m)).pcg(b, rel_tol=1e-12)
    assert np.linalg.norm(x - xo) / np.linalg.norm(xo) <= 1e-8
    log(f"PCG iterations: Jacobi V-cycle {ij.iterations}, multicolour SGS V-cycle "
        f"{im.iterations}")
    assert im.iterations <= ij.iterations
    with pytest.raises(lal.LalError):
        H32, _ = _hier(h, rp, ci, v, m, build_f32=1)
        H32.pcg_reliable(dev(b), X, rel_tol=1e-10, smoother="mcsgs")


def test_multicolour_vcycle_time_128(h):
    """One V(1,1) cycle with multicolour Gauss-Seidel at 3D Poisson 128^3 (the
    colourings built beforehand), as amg_solve runs it: the cycle (captured as
    a HIP graph) plus the residual norm that the host reads after it, the best
    of three runs of ten cycles.  At most 2 ms forward (1.91 ms measured on
    MI355X); the symmetric cycle sweeps twice per level and has twice the
    bar."""
    rp, ci, v, m = pb.poisson3d(128)
    H, _ = _hier(h, rp, ci, v, m)
    b = dev(pb.uniform(m, seed=3))
    X = torch.zeros(m, dtype=torch.float64, device="cuda")
    for sm in ("mcgs", "mcsgs"):
        H.amg_solve(b, X, 1, 1, smoother=sm, rel_tol=1e-300, abs_tol=1e-300, max_iter=0)
        torch.cuda.synchronize()
        reps, best = 10, float("inf")
        for _ in range(3):
            X.zero_()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            H.amg_solve(b, X, 1, 1, smoother=sm, rel_tol=1e-300, abs_tol=1e-300,
                        max_iter=reps - 1)
            torch.cuda.synchronize()
            best = min(best, (time.perf_counter() - t0) / reps)
        log(f"128^3 V-cycle with {sm}: {best * 1e3:.3f} ms (incl. the residual norm per cycle)")
        assert best <= (2e-3 if sm == "mcgs" else 4e-3), best


def test_multicolour_cycle_graph_replays_match_eager(h):
    """A multicolour cycle runs eagerly on its first use, is captured into a
    HIP graph on its second and replayed afterwards: every replay equals the
    eager cycle bit for bit, with unrelated eager work and synchronisations
    between the replays, for both smoothers and for the PCG preconditioner."""
    rp, ci, v, m = pb.poisson3d(32)
    H, _ = _hier(h, rp, ci, v, m)
    b = dev(pb.uniform(m, seed=8))
    X = torch.zeros(m, dtype=torch.float64, device="cuda")
    for sm in ("mcgs", "mcsgs"):
        ref = None
        for k in range(5):
            X.zero_()
            H.amg_solve(b, X, 1, 1, smoother=sm, rel_tol=1e-300, abs_tol=1e-300, max_iter=0)
            cur = host(X)
            if ref is None:
                ref = cur  # the eager cycle
            else:
                assert bits_equal(cur, ref), (sm, k)
            Y = torch.randn(m, dtype=torch.float64, device="cuda")
            assert np.isfinite((Y * 2.0).sum().item())
            torch.cuda.synchronize()
    # PCG with the symmetric multicolour V-cycle: the first solve captures the
    # preconditioner's graph after its first application; repeated solves
    # replay it and must agree with each other bit for bit
    xs = []
    for k in range(3):
        X.zero_()
        info = H.pcg(b, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=200, smoother="mcsgs")
        assert info.converged
        xs.append(host(X))
        torch.cuda.synchronize()
    assert bits_equal(xs[0], xs[1]) and bits_equal(xs[1], xs[2])
