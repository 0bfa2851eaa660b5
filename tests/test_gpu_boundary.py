"""GPU checks of the C ABI entry points that let LinAlgLib's device shims bind
one to one (backend/device/*.h):

* the three PMIS steps (device_amg_aggregation.h:40-58) driven the way
  amg_aggregation.cpp:106-215 drives them give the oracle's aggregates bit
  for bit;
* the fused y = alpha x + y with <y, z> equals lal_axpy_f64 then lal_dot_f64
  bit for bit;
* the FP32 residual and transpose agree with their FP64 counterparts.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import grid, random_csr
from tests.gpu_util import DevCSR, bits_equal, dev, empty, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def pmis_driver(h, A, d, conn):
    """compute_aggregates_using_pmis (amg_aggregation.cpp:106-215) over the
    per-step entry points: init, at most 20 rounds on a snapshot of max_state,
    the exclusive scan of the root flags, two join passes."""
    m, nnz = A.m, A.nnz
    state = empty(m, torch.int32)
    max_state = empty(m, torch.int32)
    hsh = empty(m, torch.int32)
    agg = torch.zeros(m, dtype=torch.int64, device="cuda")
    lal.call("lal_pmis_init", h.h, m, nnz, lal.ptr(d.row_ptr), lal.ptr(conn),
             lal.ptr(max_state), lal.ptr(hsh))
    rounds = 0
    for _ in range(20):
        state.copy_(max_state)
        done = C.c_int(0)
        lal.call("lal_pmis_round", h.h, m, nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                 lal.ptr(conn), lal.ptr(state), lal.ptr(hsh), lal.ptr(agg), lal.ptr(max_state),
                 C.byref(done))
        rounds += 1
        if done.value:
            break
    lal.call("lal_exclusive_scan_i64", h.h, m, lal.ptr(agg))
    for _ in range(2):
        state.copy_(max_state)
        lal.call("lal_pmis_join", h.h, m, nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                 lal.ptr(conn), lal.ptr(state), lal.ptr(agg), lal.ptr(max_state))
    return host(agg), rounds


@pytest.mark.parametrize("name", ["diff12", "3d8", "2d128"])
def test_pmis_steps_match_oracle(h, oracle, name):
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    d = DevCSR(A)
    for eps in (0.08, 0.25):
        conn = empty(A.nnz, torch.int32)
        lal.call("lal_sa_strength_f64", h.h, m, A.nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                 lal.ptr(d.val), eps, lal.ptr(conn))
        c_ref = oracle.strength(A, eps)
        assert np.array_equal(host(conn)[: A.nnz], c_ref)
        agg, rounds = pmis_driver(h, A, d, conn)
        log(f"pmis steps {name} eps={eps}: {rounds} rounds")
        assert np.array_equal(agg, oracle.aggregates(A, c_ref)), (name, eps)
        # the fused entry point gives the same aggregates
        agg2 = empty(m, torch.int64)
        lal.call("lal_pmis_aggregate", h.h, m, A.nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                 lal.ptr(conn), lal.ptr(agg2))
        assert np.array_equal(host(agg2), agg)


@pytest.mark.parametrize("n", [1, 700, 1024, 1025, 100_003, 3_000_000])
def test_axpy_dot_fused_equals_separate(h, n):
    rng = np.random.default_rng(n)
    x, y, z = (rng.uniform(-1, 1, n) for _ in range(3))
    for alpha in (1.0, -0.37):
        for z_is_y in (False, True):
            X, Y1, Y2, Z = dev(x), dev(y), dev(y), dev(z)
            r1, r2 = C.c_double(), C.c_double()
            lal.call("lal_axpy_f64", h.h, n, alpha, lal.ptr(X), lal.ptr(Y1))
            lal.call("lal_dot_f64", h.h, n, lal.ptr(Y1), lal.ptr(Y1 if z_is_y else Z), C.byref(r1))
            lal.call("lal_axpy_dot_f64", h.h, n, alpha, lal.ptr(X), lal.ptr(Y2),
                     lal.ptr(Y2 if z_is_y else Z), C.byref(r2))
            assert bits_equal(host(Y1), host(Y2)), (n, alpha)
            assert bits_equal(np.array([r1.value]), np.array([r2.value])), (n, alpha, z_is_y)
            yn = alpha * x + y
            ref = yn @ (yn if z_is_y else z)
            assert abs(r2.value - ref) <= 1e-12 * (np.abs(yn) @ np.abs(yn if z_is_y else z))


def test_residual_f32_and_transpose_f32(h):
    rp, ci, v, m = grid("diff12")
    A = CSR(rp, ci, v, m, m)
    d = DevCSR(A)
    d32 = DevCSR(A, f32=True)
    desc = lal.CsrDescr(h, m, m, A.nnz, d.row_ptr, d.col_ind)
    rng = np.random.default_rng(3)
    x = rng.uniform(-1, 1, m).astype(np.float32)
    b = rng.uniform(-1, 1, m).astype(np.float32)
    R = empty(m, torch.float32)
    X, Bv = dev(x), dev(b)  # kept alive until the kernels have read them
    for dd in (desc, None):
        lal.call("lal_residual_f32", h.h, dd.d if dd else None, m, m, A.nnz, lal.ptr(d32.row_ptr),
                 lal.ptr(d32.col_ind), lal.ptr(d32.val), lal.ptr(X), lal.ptr(Bv), lal.ptr(R))
        r = host(R).astype(np.float64)
        v32 = v.astype(np.float32).astype(np.float64)
        ax = np.zeros(m)
        bound = np.zeros(m)
        for i in range(m):
            s = slice(rp[i], rp[i + 1])
            ax[i] = v32[s] @ x[ci[s]].astype(np.float64)
            bound[i] = np.abs(v32[s]) @ np.abs(x[ci[s]].astype(np.float64))
        ref = b.astype(np.float64) - ax
        lens = np.diff(rp)
        assert np.all(np.abs(r - ref) <= (lens + 3) * 2.0 ** -24 * (bound + np.abs(b)))
    # transpose: the FP64 pattern, the values rounded to FP32
    P = random_csr(np.random.default_rng(5), 300, 200, 12)
    dP, dP32 = DevCSR(P), DevCSR(P, f32=True)
    out = []
    for src, dt in ((dP, torch.float64), (dP32, torch.float32)):
        trp = empty(P.n + 1, torch.int32)
        tci = empty(P.nnz, torch.int32)
        tv = empty(P.nnz, dt)
        fn = "lal_transpose_f64" if dt == torch.float64 else "lal_transpose_f32"
        lal.call(fn, h.h, P.m, P.n, P.nnz, lal.ptr(src.row_ptr), lal.ptr(src.col_ind),
                 lal.ptr(src.val), lal.ptr(trp), lal.ptr(tci), lal.ptr(tv))
        out.append((host(trp), host(tci), host(tv)))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2].astype(np.float32), out[1][2])
