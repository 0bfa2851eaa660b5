"""Differentiable solve: adjoint gradients of L = 0.5 ||x - x_t||^2 with
x = A(k)^-1 b for face conductances k, against the FP64 CPU oracle and
central finite differences."""
import numpy as np
import pytest
import torch

from linalglib_amd import problems as pb
from linalglib_amd.torch_solve import saamg_solve
from oracle.oracle import CSR
from tests.gpu_util import log

pytestmark = pytest.mark.gpu


def setup_problem(nx, seed=0):
    rp, ci, m, nf, entry, face, sign = pb.diffusion_faces(nx)
    rng = np.random.default_rng(seed)
    k = np.exp(rng.standard_normal(nf))
    b = rng.standard_normal(m)
    xt = rng.standard_normal(m)
    return rp, ci, m, nf, entry, face, sign, k, b, xt


def face_cells(nx):
    """(p, q) cells of every face; q = -1 for Dirichlet faces."""
    rp, ci, m, nf, entry, face, sign = pb.diffusion_faces(nx)
    p = np.full(nf, -1)
    q = np.full(nf, -1)
    rows = np.repeat(np.arange(m), np.diff(rp))
    off = sign < 0
    e = entry[off]
    f = face[off]
    # each interior face appears twice among the off-diagonal entries
    for fi, r, c in zip(f, rows[e], ci[e]):
        if p[fi] < 0:
            p[fi], q[fi] = min(r, c), max(r, c)
    diag = ~off
    for fi, r in zip(face[diag], rows[entry[diag]]):
        if p[fi] < 0:
            p[fi] = r
    return p, q


def oracle_gradient(oracle, rp, ci, m, val, b, xt, p, q):
    A = CSR(rp, ci, val, m, m)
    h = oracle.saamg_setup(A)
    x, _, _ = h.pcg(b, rel_tol=1e-13, coarse_solver=1)
    lam, _, _ = h.pcg(x - xt, rel_tol=1e-13, coarse_solver=1)
    lq = np.where(q >= 0, lam[np.maximum(q, 0)], 0.0)
    xq = np.where(q >= 0, x[np.maximum(q, 0)], 0.0)
    dk = -(lam[p] - lq) * (x[p] - xq)  # closed form of sum_ij dL/dA_ij dA_ij/dk_f
    return x, lam, dk


def gpu_loss_and_grad(rp, ci, entry, face, sign, k, b, xt, mixed=False):
    dev = torch.device("cuda:0")
    K = torch.tensor(k, device=dev, requires_grad=True)
    B = torch.tensor(b, device=dev, requires_grad=True)
    RP = torch.from_numpy(rp).to(dev)
    CI = torch.from_numpy(ci).to(dev)
    E = torch.from_numpy(entry).to(dev)
    F = torch.from_numpy(face).to(dev)
    S = torch.from_numpy(sign).to(dev)
    val = torch.zeros(len(ci), dtype=torch.float64, device=dev).index_add(0, E, S * K[F])
    x = saamg_solve(val, B, RP, CI, mixed=mixed, tol=1e-12)
    L = 0.5 * ((x - torch.from_numpy(xt).to(dev)) ** 2).sum()
    L.backward()
    torch.cuda.synchronize()
    return L.item(), x.detach().cpu().numpy(), K.grad.cpu().numpy(), B.grad.cpu().numpy()


@pytest.mark.parametrize("mixed", [False, "ir", "vcycle", "reliable"])
def test_gradient_matches_fp64_oracle(oracle, mixed):
    nx = 20
    rp, ci, m, nf, entry, face, sign, k, b, xt = setup_problem(nx, seed=3)
    val = pb.assemble_faces(len(ci), entry, face, sign, k)
    p, q = face_cells(nx)
    log(f"adjoint vs oracle nx={nx} mixed={mixed}")
    xo, lam, dko = oracle_gradient(oracle, rp, ci, m, val, b, xt, p, q)
    L, x, dk, db = gpu_loss_and_grad(rp, ci, entry, face, sign, k, b, xt, mixed)
    assert np.linalg.norm(x - xo) <= 1e-10 * np.linalg.norm(xo)
    assert np.linalg.norm(db - lam) <= 1e-6 * np.linalg.norm(lam)
    assert np.linalg.norm(dk - dko) <= 1e-6 * np.linalg.norm(dko)
    assert np.abs(dk - dko).max() <= 1e-6 * np.abs(dko).max()


def test_gradient_central_differences(oracle):
    nx = 10
    rp, ci, m, nf, entry, face, sign, k, b, xt = setup_problem(nx, seed=5)
    log("adjoint vs central differences")
    _, _, dk, _ = gpu_loss_and_grad(rp, ci, entry, face, sign, k, b, xt)
    rng = np.random.default_rng(11)
    faces = rng.choice(nf, size=16, replace=False)
    h = 1e-6

    def loss(kk):
        val = pb.assemble_faces(len(ci), entry, face, sign, kk)
        hh = oracle.saamg_setup(CSR(rp, ci, val, m, m))
        x, _, _ = hh.pcg(b, rel_tol=1e-14, abs_tol=1e-300, coarse_solver=1)
        return 0.5 * np.sum((x - xt) ** 2)

    for f in faces:
        kp = k.copy()
        km = k.copy()
        kp[f] += h
        km[f] -= h
        fd = (loss(kp) - loss(km)) / (2 * h)
        assert abs(fd - dk[f]) <= 1e-4 * max(abs(dk[f]), 1e-3 * np.abs(dk).max()), (f, fd, dk[f])
