"""kNN graph Laplacians (symmetrised 12-NN of uniform points, L(w) + 1e-2 I):
the device-built hierarchy against the CPU oracle bit for bit, PCG and both
mixed-precision schemes against the oracle solve, and dL/dw through the
differentiable solve against the oracle's closed-form gradient
dL/dw_e = -(lambda_p - lambda_q)(x_p - x_q)."""
import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from linalglib_amd.torch_solve import saamg_solve
from oracle.oracle import CSR
from tests.gpu_util import DevCSR, dev, empty, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def knn_problem(n, seed, order="morton"):
    rp, ci, m, E, entry, edge, sign, dpos = pb.knn_laplacian(n, seed=seed, order=order)
    rng = np.random.default_rng(seed + 100)
    w = rng.uniform(0.5, 2.0, len(E))
    val = pb.assemble_knn(len(ci), entry, edge, sign, dpos, w)
    b = rng.standard_normal(m)
    return rp, ci, m, E, entry, edge, sign, dpos, w, val, b


@pytest.mark.parametrize("order", ["morton", "random"])
def test_knn_hierarchy_and_pcg(h, oracle, order):
    n = 6000
    rp, ci, m, E, entry, edge, sign, dpos, w, val, b = knn_problem(n, 4, order)
    A = CSR(rp, ci, val, m, m)
    log(f"knn hierarchy n={n} order={order}")
    d = DevCSR(A)
    H = lal.Hierarchy(h, m, A.nnz, d.row_ptr, d.col_ind, d.val, build_f32=1)
    ho = oracle.saamg_setup(A)
    assert H.levels == ho.levels
    for lvl in range(H.levels + 1):
        for kind in ("A", "P", "R") if lvl < H.levels else ("A",):
            grp, gci, gv, _, _ = H.level(lvl, kind)
            ref = ho.get(lvl, kind)
            assert np.array_equal(grp, ref.row_ptr), (lvl, kind)
            assert np.array_equal(gci, ref.col_ind), (lvl, kind)
            assert np.array_equal(gv.view(np.int64), ref.val.view(np.int64)), (lvl, kind)
    xo, ito, _ = ho.pcg(b, rel_tol=1e-10, coarse_solver=1)
    B = dev(b)
    x = empty(m, torch.float64)
    info = H.pcg(B, x, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    assert info.converged and abs(info.iterations - ito) <= 1, (info.iterations, ito)
    assert np.linalg.norm(host(x) - xo) <= 1e-9 * np.linalg.norm(xo)
    for name, solve in (("vcycle", H.pcg_mixed), ("ir", H.ir), ("reliable", H.pcg_reliable)):
        x.zero_()
        info = solve(B, x, rel_tol=1e-10, abs_tol=1e-300)
        xg = host(x)
        r = oracle.residual(A, xg, b)
        assert info.converged, name
        assert np.abs(r).max() <= 1e-9 * np.abs(b).max(), name
        assert np.linalg.norm(xg - xo) <= 1e-7 * np.linalg.norm(xo), name


@pytest.mark.parametrize("mixed", [False, "vcycle", "reliable"])
def test_knn_edge_weight_gradient(oracle, mixed):
    n = 4000
    rp, ci, m, E, entry, edge, sign, dpos, w, val, b = knn_problem(n, 7)
    xt = np.random.default_rng(9).standard_normal(m)
    log(f"knn gradient n={n} mixed={mixed}")
    ho = oracle.saamg_setup(CSR(rp, ci, val, m, m))
    xo, _, _ = ho.pcg(b, rel_tol=1e-13, coarse_solver=1)
    lam, _, _ = ho.pcg(xo - xt, rel_tol=1e-13, coarse_solver=1)
    p, q = E[:, 0], E[:, 1]
    dwo = -(lam[p] - lam[q]) * (xo[p] - xo[q])

    cuda = torch.device("cuda:0")
    W = torch.tensor(w, device=cuda, requires_grad=True)
    Ent = torch.from_numpy(entry).to(cuda)
    Edg = torch.from_numpy(edge).to(cuda)
    S = torch.from_numpy(sign).to(cuda)
    D = torch.from_numpy(dpos).to(cuda)
    V = torch.zeros(len(ci), dtype=torch.float64, device=cuda).index_add(0, Ent, S * W[Edg])
    V = V.index_add(0, D, torch.full((m,), 1e-2, dtype=torch.float64, device=cuda))
    x = saamg_solve(V, torch.from_numpy(b).to(cuda), torch.from_numpy(rp).to(cuda),
                    torch.from_numpy(ci).to(cuda), mixed=mixed, tol=1e-12)
    L = 0.5 * ((x - torch.from_numpy(xt).to(cuda)) ** 2).sum()
    L.backward()
    dw = W.grad.cpu().numpy()
    xg = x.detach().cpu().numpy()
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo)
    assert np.linalg.norm(dw - dwo) <= 1e-6 * np.linalg.norm(dwo)
