"""GPU parity of the BLAS-1 and CSR SpMV kernels against the CPU oracle,
through the C ABI."""
import ctypes as C

import numpy as np
import pytest
import torch

from linalglib_amd import lal, problems as pb
from oracle.oracle import CSR
from tests.golden.make_golden import random_csr
from tests.gpu_util import DevCSR, bits_equal, dev, empty, host, log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def h():
    return lal.Handle()


def irregular_csr(seed=0):
    """Empty rows, short rows and rows longer than one LDS chunk (2048)."""
    rng = np.random.default_rng(seed)
    m, n = 700, 5000
    lens = rng.integers(0, 12, m)
    lens[::97] = 0
    lens[5] = 3000
    lens[300] = 2049
    lens[301] = 2048
    lens[640:648] = 700
    rows = [np.sort(rng.choice(n, size=k, replace=False)) for k in lens]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    ci = np.concatenate(rows).astype(np.int32)
    return CSR(rp, ci, rng.uniform(-1, 1, rp[-1]), m, n)


def matrices():
    rp, ci, v, m = pb.poisson3d(16)
    yield "poisson3d16", CSR(rp, ci, v, m, m)
    yield "irregular", irregular_csr()
    yield "random", random_csr(np.random.default_rng(5), 1000, 800, 40)


@pytest.mark.parametrize("exact", [1, 0])
def test_csrmv_f64(h, oracle, exact):
    for name, A in matrices():
        log(f"csrmv_f64 {name} exact={exact}")
        d = DevCSR(A)
        desc = lal.CsrDescr(h, A.m, A.n, A.nnz, d.row_ptr, d.col_ind, exact=exact)
        rng = np.random.default_rng(1)
        x = rng.uniform(-1, 1, A.n)
        y0 = rng.uniform(-1, 1, A.m)
        X = dev(x)
        for alpha, beta in ((1.0, 0.0), (1.0, 1.0), (-2.0, 0.5)):
            y = dev(y0)
            lal.call("lal_csrmv_f64", h.h, desc.d, A.m, A.n, A.nnz, alpha, lal.ptr(d.row_ptr),
                     lal.ptr(d.col_ind), lal.ptr(d.val), lal.ptr(X), beta, lal.ptr(y))
            got = host(y)
            ref = oracle.csrmv(A, x, alpha, beta, y0)
            if exact:
                assert bits_equal(got, ref), (name, alpha, beta)
            else:
                # rounded products summed in another order: the recursive
                # summation bound (len + 2) u sum|a_ij x_j| (u = 2^-53)
                absr = oracle.csrmv(CSR(A.row_ptr, A.col_ind, np.abs(A.val), A.m, A.n),
                                    np.abs(x)) * abs(alpha) + abs(beta * y0)
                lens = np.diff(A.row_ptr)
                tol = (lens + 3) * 1.12e-16 * absr + 1e-300
                assert np.all(np.abs(got - ref) <= tol), (name, alpha, beta)
        # without a descriptor: 256 rows per block
        y = empty(A.m, d.dtype)
        lal.call("lal_csrmv_f64", h.h, None, A.m, A.n, A.nnz, 1.0, lal.ptr(d.row_ptr),
                 lal.ptr(d.col_ind), lal.ptr(d.val), lal.ptr(X), 0.0, lal.ptr(y))
        assert bits_equal(host(y)[: A.m], oracle.csrmv(A, x))


@pytest.mark.parametrize("exact", [1, 0, None])
def test_csrmv_empty_inputs(h, oracle, exact):
    """Degenerate sizes: no rows (a no-op), and rows without nonzeros
    (y = beta y, the residual returns b), with an exact or fast descriptor or
    none."""
    log(f"csrmv empty inputs exact={exact}")
    none = CSR(np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), 0, 5)
    d0 = DevCSR(none)
    desc0 = None if exact is None else lal.CsrDescr(h, 0, 5, 0, d0.row_ptr, d0.col_ind,
                                                    exact=exact)
    lal.call("lal_csrmv_f64", h.h, desc0.d if desc0 else None, 0, 5, 0, 1.0,
             lal.ptr(d0.row_ptr), None, None, None, 0.0, lal.ptr(empty(1, torch.float64)))
    m, n = 300, 40
    A = CSR(np.zeros(m + 1, np.int32), np.zeros(0, np.int32), np.zeros(0), m, n)
    d = DevCSR(A)
    desc = None if exact is None else lal.CsrDescr(h, m, n, 0, d.row_ptr, d.col_ind, exact=exact)
    x = np.random.default_rng(2).uniform(-1, 1, n)
    y0 = np.random.default_rng(3).uniform(-1, 1, m)
    X = dev(x)
    for alpha, beta in ((1.0, 0.0), (2.0, -0.5)):
        y = dev(y0)
        lal.call("lal_csrmv_f64", h.h, desc.d if desc else None, m, n, 0, alpha,
                 lal.ptr(d.row_ptr), None, None, lal.ptr(X), beta, lal.ptr(y))
        assert bits_equal(host(y), oracle.csrmv(A, x, alpha, beta, y0)), (alpha, beta)
    r = empty(m, torch.float64)
    B = dev(y0)
    lal.call("lal_residual_f64", h.h, desc.d if desc else None, m, n, 0, lal.ptr(d.row_ptr),
             None, None, lal.ptr(X), lal.ptr(B), lal.ptr(r))
    assert bits_equal(host(r), y0)


def test_csrmv_f32(h, oracle):
    for name, A in matrices():
        log(f"csrmv_f32 {name}")
        d = DevCSR(A, f32=True)
        desc = lal.CsrDescr(h, A.m, A.n, A.nnz, d.row_ptr, d.col_ind)
        x = np.random.default_rng(2).uniform(-1, 1, A.n).astype(np.float32)
        y = empty(A.m, d.dtype)
        X = dev(x)
        lal.call("lal_csrmv_f32", h.h, desc.d, A.m, A.n, A.nnz, 1.0, lal.ptr(d.row_ptr),
                 lal.ptr(d.col_ind), lal.ptr(d.val), lal.ptr(X), 0.0, lal.ptr(y))
        A32 = CSR(A.row_ptr, A.col_ind, A.val.astype(np.float32).astype(np.float64), A.m, A.n)
        ref = oracle.csrmv(A32, x.astype(np.float64))
        absr = oracle.csrmv(CSR(A.row_ptr, A.col_ind, np.abs(A32.val), A.m, A.n),
                            np.abs(x.astype(np.float64)))
        got = host(y)[: A.m].astype(np.float64)
        # FP32 accumulation: rows of up to 3000 terms, 2e-7 per term bound
        lens = np.diff(A.row_ptr)
        assert np.all(np.abs(got - ref) <= 1.2e-7 * (lens + 1) * absr + 1e-30), name


def _fast_tol(A, x, alpha, beta, y0, oracle, u):
    absr = oracle.csrmv(CSR(A.row_ptr, A.col_ind, np.abs(A.val), A.m, A.n),
                        np.abs(x)) * abs(alpha) + abs(beta * y0)
    return (np.diff(A.row_ptr) + 3) * u * absr + 1e-300


@pytest.mark.parametrize("f32", [False, True])
def test_csrmv_fast_tails_and_alignment(h, oracle, f32):
    """The fast kernel loads four nonzeros per lane from 16-byte aligned
    arrays: matrices whose nnz is not a multiple of four (and tiny ones) end in
    a scalar tail; column / value arrays at an odd element offset take the
    product-staging kernel.  Both must agree with the oracle to rounding."""
    import torch

    rng = np.random.default_rng(9)
    cases = []
    for nnz_target in (1, 2, 3, 5, 6, 7, 4097, 4098, 4099):
        m = max(1, nnz_target // 3)
        lens = np.zeros(m, np.int64)
        lens[:] = nnz_target // m
        lens[: nnz_target - lens.sum()] += 1
        n = 50
        rows = [np.sort(rng.choice(n, size=int(k), replace=False)) for k in lens]
        rp = np.zeros(m + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        cases.append(CSR(rp, np.concatenate(rows).astype(np.int32),
                         rng.uniform(-1, 1, rp[-1]), m, n))
    cases.append(irregular_csr(3))
    vdt = torch.float32 if f32 else torch.float64
    fn = "lal_csrmv_f32" if f32 else "lal_csrmv_f64"
    u = 1.2e-7 if f32 else 1.12e-16
    for A in cases:
        for offset in (0, 1):
            log(f"csrmv fast f32={f32} nnz={A.nnz} offset={offset}")
            RP = dev(A.row_ptr, np.int32)
            CIb = torch.zeros(A.nnz + 1, dtype=torch.int32, device="cuda")
            Vb = torch.zeros(A.nnz + 1, dtype=vdt, device="cuda")
            CIb[offset: offset + A.nnz] = dev(A.col_ind, np.int32)
            Vb[offset: offset + A.nnz] = dev(A.val).to(vdt)
            CI, V = CIb[offset:], Vb[offset:]
            desc = lal.CsrDescr(h, A.m, A.n, A.nnz, RP, CI)
            x = rng.uniform(-1, 1, A.n)
            X = dev(x).to(vdt)
            y = empty(A.m, vdt)
            lal.call(fn, h.h, desc.d, A.m, A.n, A.nnz, 1.0, lal.ptr(RP), lal.ptr(CI),
                     lal.ptr(V), lal.ptr(X), 0.0, lal.ptr(y))
            Ar = A
            xr = x
            if f32:
                Ar = CSR(A.row_ptr, A.col_ind, A.val.astype(np.float32).astype(np.float64),
                         A.m, A.n)
                xr = x.astype(np.float32).astype(np.float64)
            ref = oracle.csrmv(Ar, xr)
            got = host(y)[: A.m].astype(np.float64)
            tol = _fast_tol(Ar, xr, 1.0, 0.0, np.zeros(A.m), oracle, u)
            assert np.all(np.abs(got - ref) <= tol), (A.nnz, offset)


def test_residual_and_diagonal(h, oracle):
    for name, A in matrices():
        if A.m != A.n:
            continue
        log(f"residual {name}")
        d = DevCSR(A)
        desc = lal.CsrDescr(h, A.m, A.n, A.nnz, d.row_ptr, d.col_ind, exact=True)
        rng = np.random.default_rng(3)
        x, b = rng.uniform(-1, 1, A.n), rng.uniform(-1, 1, A.m)
        r = empty(A.m, d.dtype)
        X, B = dev(x), dev(b)
        lal.call("lal_residual_f64", h.h, desc.d, A.m, A.n, A.nnz, lal.ptr(d.row_ptr),
                 lal.ptr(d.col_ind), lal.ptr(d.val), lal.ptr(X), lal.ptr(B), lal.ptr(r))
        assert bits_equal(host(r), oracle.residual(A, x, b))
        dg = dev(np.full(A.m, 7.0))
        lal.call("lal_diagonal_f64", h.h, A.m, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                 lal.ptr(d.val), lal.ptr(dg))
        ref = np.full(A.m, 7.0)
        oracle.L.or_diagonal(A.m, A.row_ptr, A.col_ind, A.val, ref)
        assert bits_equal(host(dg), ref)


def test_blas1(h, oracle):
    for n in (1, 255, 257, 100003, 3_000_000):
        log(f"blas1 n={n}")
        rng = np.random.default_rng(n)
        x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
        X, Y = dev(x), dev(y)
        for alpha in (1.0, -0.75):
            Z = Y.clone()
            lal.call("lal_axpy_f64", h.h, n, alpha, lal.ptr(X), lal.ptr(Z))
            assert bits_equal(host(Z), oracle.axpy(alpha, x, y))
        for a, b in ((0.0, 2.0), (3.0, 0.0), (1.0, -0.5)):
            Z = Y.clone()
            lal.call("lal_axpby_f64", h.h, n, a, lal.ptr(X), b, lal.ptr(Z))
            assert bits_equal(host(Z), oracle.axpby(a, x, b, y))
        r1, r2 = C.c_double(), C.c_double()
        lal.call("lal_dot_f64", h.h, n, lal.ptr(X), lal.ptr(Y), C.byref(r1))
        lal.call("lal_dot_f64", h.h, n, lal.ptr(X), lal.ptr(Y), C.byref(r2))
        assert bits_equal(r1.value, r2.value)  # deterministic
        ref = oracle.dot(x, y)
        assert abs(r1.value - ref) <= 1e-15 * np.dot(np.abs(x), np.abs(y)) * 4 * np.log2(n + 2)
        nrm = C.c_double()
        lal.call("lal_norm_inf_f64", h.h, n, lal.ptr(X), C.byref(nrm))
        assert nrm.value == oracle.norm_inf(x)
        # closed form of the reference's dot test: 2 . 3 = 6 n
        two, three = dev(np.full(n, 2.0)), dev(np.full(n, 3.0))
        lal.call("lal_dot_f64", h.h, n, lal.ptr(two), lal.ptr(three), C.byref(r1))
        assert r1.value == 6.0 * n


def test_exclusive_scan(h):
    for n in (1, 2, 2047, 2048, 2049, 1_000_003, 5_000_000):
        log(f"scan n={n}")
        rng = np.random.default_rng(n)
        a = rng.integers(0, 3, n)
        for dt, fn in ((np.int32, "lal_exclusive_scan_i32"), (np.int64, "lal_exclusive_scan_i64")):
            t = dev(a, dt)
            lal.call(fn, h.h, n, lal.ptr(t))
            ref = np.concatenate([[0], np.cumsum(a)[:-1]]).astype(dt)
            assert np.array_equal(host(t), ref)


def test_invalid_pattern_refused(h):
    A = random_csr(np.random.default_rng(9), 50, 40, 5)
    ci = A.col_ind.copy()
    ci[len(ci) // 2] = 40  # out of range
    d = DevCSR(CSR(A.row_ptr, ci, A.val, 50, 40))
    with pytest.raises(lal.LalError) as e:
        lal.CsrDescr(h, 50, 40, A.nnz, d.row_ptr, d.col_ind)
    assert e.value.status == 1
