"""GPU parity of the SA-AMG setup, the V-cycle and the PCG / mixed-precision
solves against the CPU oracle and the reference's golden fixtures, through the
C ABI."""
import ctypes as C
import hashlib

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


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def csr_equal(got, ref):
    rp, ci, v, m, n = got
    return (m == ref.m and n == ref.n and np.array_equal(rp, ref.row_ptr)
            and np.array_equal(ci, ref.col_ind) and bits_equal(v, ref.val))


# ---------------------------------------------------------------- setup pieces


def test_strength_aggregates_prolongator(h, oracle):
    cases = [("diff12", grid("diff12")), ("poisson3d10", pb.poisson3d(10)),
             ("poisson2d33", pb.poisson2d(33))]
    # the oracle's own Galerkin operators of 3D Poisson 24^3: 36 and 90
    # entries per row, which the setup passes walk with 16 and 64 lanes per
    # row instead of one
    rp0, ci0, v0, m0 = pb.poisson3d(24)
    H = oracle.saamg_setup(CSR(rp0, ci0, v0, m0, m0))
    for lvl in (1, 2):
        Ag = H.get(lvl, "A")
        cases.append((f"galerkin{lvl}", (Ag.row_ptr, Ag.col_ind, Ag.val, Ag.m)))
    for name, (rp, ci, v, m) in cases:
        A = CSR(rp, ci, v, m, m)
        d = DevCSR(A)
        for eps in (0.08, 0.25, 0.01):
            log(f"setup pieces {name} eps={eps}")
            conn = empty(A.nnz, torch.int32)
            lal.call("lal_sa_strength_f64", h.h, m, A.nnz, lal.ptr(d.row_ptr),
                     lal.ptr(d.col_ind), lal.ptr(d.val), eps, lal.ptr(conn))
            c_ref = oracle.strength(A, eps)
            assert np.array_equal(host(conn)[: A.nnz], c_ref)
            agg = empty(m, torch.int64)
            lal.call("lal_pmis_aggregate", h.h, m, A.nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
                     lal.ptr(conn), lal.ptr(agg))
            a_ref = oracle.aggregates(A, c_ref)
            assert np.array_equal(host(agg), a_ref)
            prp = empty(m + 1, torch.int32)
            nnz_p, nc = C.c_int(), C.c_int()
            lal.call("lal_sa_prolongator_nnz", h.h, m, A.nnz, lal.ptr(d.row_ptr),
                     lal.ptr(d.col_ind), lal.ptr(conn), lal.ptr(agg), lal.ptr(prp),
                     C.byref(nnz_p), C.byref(nc))
            P_ref = oracle.prolongator(A, c_ref, a_ref)
            assert (nnz_p.value, nc.value) == (P_ref.nnz, P_ref.n)
            pci = empty(nnz_p.value, torch.int32)
            pv = empty(nnz_p.value, torch.float64)
            lal.call("lal_sa_prolongator_fill_f64", h.h, m, A.nnz, lal.ptr(d.row_ptr),
                     lal.ptr(d.col_ind), lal.ptr(d.val), lal.ptr(conn), lal.ptr(agg), 2.0 / 3.0,
                     lal.ptr(prp), lal.ptr(pci), lal.ptr(pv))
            got = (host(prp), host(pci)[: P_ref.nnz], host(pv)[: P_ref.nnz], m, nc.value)
            assert csr_equal(got, P_ref)


def _gpu_transpose(h, A):


    d = DevCSR(A)
    trp = empty(A.n + 1, torch.int32)
    tci = empty(A.nnz, torch.int32)
    tv = empty(A.nnz, torch.float64)
    lal.call("lal_transpose_f64", h.h, A.m, A.n, A.nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind),
             lal.ptr(d.val), lal.ptr(trp), lal.ptr(tci), lal.ptr(tv))
    return host(trp), host(tci)[: A.nnz], host(tv)[: A.nnz], A.n, A.m


def _gpu_spgemm(h, A, B):


    da, db = DevCSR(A), DevCSR(B)
    crp = empty(A.m + 1, torch.int32)
    nnz = C.c_int()
    lal.call("lal_csrgemm_nnz", h.h, A.m, A.n, B.n, A.nnz, lal.ptr(da.row_ptr),
             lal.ptr(da.col_ind), B.nnz, lal.ptr(db.row_ptr), lal.ptr(db.col_ind), lal.ptr(crp),
             C.byref(nnz))
    cci = empty(nnz.value, torch.int32)
    cv = empty(nnz.value, torch.float64)
    lal.call("lal_csrgemm_solve_f64", h.h, A.m, A.n, B.n, A.nnz, lal.ptr(da.row_ptr),
             lal.ptr(da.col_ind), lal.ptr(da.val), B.nnz, lal.ptr(db.row_ptr),
             lal.ptr(db.col_ind), lal.ptr(db.val), lal.ptr(crp), lal.ptr(cci), lal.ptr(cv))
    return host(crp), host(cci)[: nnz.value], host(cv)[: nnz.value], A.m, B.n


def test_transpose_and_spgemm_golden(h, oracle, golden):
    g = golden("ops")

    def mk(p, m, n):
        return CSR(g[p + "_rp"], g[p + "_ci"], g[p + "_v"], m, n)

    log("spgemm golden")
    for a, b, c, (m, k, n) in (("g1", "g2", "g", (120, 90, 110)), ("h1", "h2", "h", (80, 70, 60))):
        got = _gpu_spgemm(h, mk(a, m, k), mk(b, k, n))
        assert csr_equal(got, mk(c, m, n)), c
    got = _gpu_transpose(h, mk("h1", 80, 70))
    assert csr_equal(got, mk("t", 70, 80))


def test_transpose_row_length_classes(h, oracle):
    """Transposed rows of every length the sort distinguishes: up to 64
    entries (one per lane), up to 256 (several per lane) and beyond 256 (the
    insertion-sort path), against the oracle's stable transpose."""
    rng = np.random.default_rng(5)
    m, n = 3000, 40
    p = (np.arange(n) + 1.0) ** 2
    p /= p.sum()
    rows = [np.sort(rng.choice(n, size=int(rng.integers(0, 6)), replace=False, p=p))
            for _ in range(m)]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    T = CSR(rp, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, rp[-1]), m, n)
    counts = np.bincount(T.col_ind, minlength=n)
    assert counts.min() <= 64 and ((counts > 64) & (counts <= 256)).any() and counts.max() > 256
    log("transpose row-length classes")
    assert csr_equal(_gpu_transpose(h, T), oracle.transpose(T))


def test_spgemm_bins_and_global_tables(h, oracle):
    """Rows spanning every hash-table size, including rows beyond the LDS
    tables (products > 2048) and empty rows."""
    rng = np.random.default_rng(21)
    m, k, n = 300, 400, 6000
    lens = rng.integers(0, 6, m)
    lens[:40] = rng.integers(40, 120, 40)  # many products
    rows = [np.sort(rng.choice(k, size=x, replace=False)) for x in lens]
    rp = np.zeros(m + 1, np.int32)
    rp[1:] = np.cumsum(lens)
    A = CSR(rp, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, rp[-1]), m, k)
    B = random_csr(rng, k, n, 90)
    log("spgemm bins")
    got = _gpu_spgemm(h, A, B)
    assert csr_equal(got, oracle.csrgemm(A, B))
    # a narrow B (few columns) caps the tables by the column count
    Bn = random_csr(rng, k, 50, 30)
    assert csr_equal(_gpu_spgemm(h, A, Bn), oracle.csrgemm(A, Bn))
    T = random_csr(rng, 500, 3000, 40)
    assert csr_equal(_gpu_transpose(h, T), oracle.transpose(T))
    # symbolic tables beyond the largest LDS table (> 16384 products over more
    # than 16384 columns) take global memory; the numeric pass sizes the same
    # rows by their nonzero count
    k2, n2 = 600, 60000
    lens2 = np.full(30, 3)
    lens2[:6] = 250
    rows2 = [np.sort(rng.choice(k2, size=x, replace=False)) for x in lens2]
    rp2 = np.zeros(31, np.int32)
    rp2[1:] = np.cumsum(lens2)
    A2 = CSR(rp2, np.concatenate(rows2).astype(np.int32), rng.uniform(-1, 1, rp2[-1]), 30, k2)
    B2 = random_csr(rng, k2, n2, 100)
    assert csr_equal(_gpu_spgemm(h, A2, B2), oracle.csrgemm(A2, B2))


def test_spgemm_lane_groups_with_long_rows(h, oracle):
    """Rows of C whose tables are small but whose A and B rows are long: the
    lane groups (8 lanes for 64-slot tables, 16 for 128) then take several
    chunks of the A row and several passes over each B row.  C has 12 or 48
    columns, so the tables stay at 64 or 128 slots in both passes, while A
    rows hold 20-40 entries and B rows 10 or 20 (more than the group's 8 or
    16 lanes); bit for bit against the oracle."""
    rng = np.random.default_rng(23)
    for n, blen in ((12, 10), (48, 20)):
        m, k = 500, 300
        lens = rng.integers(20, 41, m)
        rows = [np.sort(rng.choice(k, size=x, replace=False)) for x in lens]
        rp = np.zeros(m + 1, np.int32)
        rp[1:] = np.cumsum(lens)
        A = CSR(rp, np.concatenate(rows).astype(np.int32), rng.uniform(-1, 1, rp[-1]), m, k)
        brows = [np.sort(rng.choice(n, size=blen, replace=False)) for _ in range(k)]
        brp = np.arange(k + 1, dtype=np.int32) * blen
        B = CSR(brp, np.concatenate(brows).astype(np.int32), rng.uniform(-1, 1, k * blen), k, n)
        log(f"spgemm lane groups: {n} columns, B rows of {blen}")
        assert csr_equal(_gpu_spgemm(h, A, B), oracle.csrgemm(A, B))


def test_spgemm_refuses_inconsistent_row_pointer(h, oracle):
    """The numeric pass sizes its tables by the row pointer of C; one that
    does not match the product's pattern (an entry moved between two rows, or
    an empty row that is not) is refused, and nothing outside C's rows is
    written."""
    rng = np.random.default_rng(22)
    A = random_csr(rng, 200, 150, 8)
    B = random_csr(rng, 150, 180, 9)
    da, db = DevCSR(A), DevCSR(B)
    crp = empty(A.m + 1, torch.int32)
    nnz = C.c_int()
    lal.call("lal_csrgemm_nnz", h.h, A.m, A.n, B.n, A.nnz, lal.ptr(da.row_ptr),
             lal.ptr(da.col_ind), B.nnz, lal.ptr(db.row_ptr), lal.ptr(db.col_ind), lal.ptr(crp),
             C.byref(nnz))
    good = host(crp).copy()
    lens = np.diff(good)
    r = int(np.argmax(lens > 2))
    for shift in ((r, -1, r + 1, +1), (r, +1, r + 1, -1)):
        bad = lens.copy()
        bad[shift[0]] += shift[1]
        bad[shift[2]] += shift[3]
        brp = np.zeros(A.m + 1, np.int32)
        brp[1:] = np.cumsum(bad)
        log(f"spgemm inconsistent row pointer {shift}")
        guard = empty(nnz.value + 64, torch.float64)
        gci = empty(nnz.value + 64, torch.int32)
        guard.fill_(7.0)
        gci.fill_(-7)
        bad_rp = dev(brp, np.int32)  # kept alive until the kernels have read it
        with pytest.raises(lal.LalError):
            lal.call("lal_csrgemm_solve_f64", h.h, A.m, A.n, B.n, A.nnz, lal.ptr(da.row_ptr),
                     lal.ptr(da.col_ind), lal.ptr(da.val), B.nnz, lal.ptr(db.row_ptr),
                     lal.ptr(db.col_ind), lal.ptr(db.val), lal.ptr(bad_rp),
                     lal.ptr(gci), lal.ptr(guard))
        assert np.all(host(guard)[nnz.value:] == 7.0) and np.all(host(gci)[nnz.value:] == -7)


def test_spgemm_refuses_repeated_column_in_b(h):
    """A row of B that repeats a column index would make two lanes of the hash
    pass fold into one slot at once; both passes refuse such a B
    (LAL_ERR_UNSUPPORTED) before any product is formed, and the same operands
    without the repeat multiply."""
    rng = np.random.default_rng(23)
    A = random_csr(rng, 40, 30, 5)
    B = random_csr(rng, 30, 35, 6)
    ci = B.col_ind.copy()
    r = int(np.argmax(np.diff(B.row_ptr) >= 2))
    ci[B.row_ptr[r] + 1] = ci[B.row_ptr[r]]  # row r names its first column twice
    Bd = CSR(B.row_ptr, ci, B.val, B.m, B.n)
    for Bx, refused in ((Bd, True), (B, False)):
        da, db = DevCSR(A), DevCSR(Bx)
        crp = empty(A.m + 1, torch.int32)
        nnz = C.c_int()
        args = (h.h, A.m, A.n, Bx.n, A.nnz, lal.ptr(da.row_ptr), lal.ptr(da.col_ind), Bx.nnz,
                lal.ptr(db.row_ptr), lal.ptr(db.col_ind), lal.ptr(crp), C.byref(nnz))
        if refused:
            with pytest.raises(lal.LalError) as e:
                lal.call("lal_csrgemm_nnz", *args)
            assert e.value.status == 7, e.value.status  # LAL_ERR_UNSUPPORTED
            cci = empty(64, torch.int32)
            cv = empty(64, torch.float64)
            with pytest.raises(lal.LalError):
                lal.call("lal_csrgemm_solve_f64", h.h, A.m, A.n, Bx.n, A.nnz,
                         lal.ptr(da.row_ptr), lal.ptr(da.col_ind), lal.ptr(da.val), Bx.nnz,
                         lal.ptr(db.row_ptr), lal.ptr(db.col_ind), lal.ptr(db.val),
                         lal.ptr(crp), lal.ptr(cci), lal.ptr(cv))
        else:
            lal.call("lal_csrgemm_nnz", *args)
            assert nnz.value > 0


# ------------------------------------------------------------------ hierarchy


def _hier(h, A, **kw):
    d = DevCSR(A)
    H = lal.Hierarchy(h, A.m, A.nnz, d.row_ptr, d.col_ind, d.val, **kw)
    H._keep = d
    return H


@pytest.mark.parametrize("m", [1, 100])
def test_hierarchy_without_coarse_levels(h, oracle, m):
    """A diagonal operator has no strong couplings: PMIS aggregates nothing,
    P would have no columns and the setup stops at the finest level, as the
    oracle's does; the V-cycle is then the coarsest solve alone (1x1 divide
    or dense inverse) and PCG stops at its first convergence test."""
    rng = np.random.default_rng(0)
    d = rng.uniform(1, 5, m)
    A = CSR(np.arange(m + 1, dtype=np.int32), np.arange(m, dtype=np.int32), d, m, m)
    b = rng.standard_normal(m)
    ho = oracle.saamg_setup(A)
    xo, ito, _ = ho.pcg(b, rel_tol=1e-10, coarse_solver=1)
    log(f"hierarchy without coarse levels m={m}")
    for kw in ({"exact": 1}, {"exact": 0, "build_f32": 1}):
        H = _hier(h, A, **kw)
        assert H.levels == ho.levels == 0
        x = empty(m, torch.float64)
        info = H.pcg(dev(b), x, rel_tol=1e-10, abs_tol=1e-300, max_iter=50)
        assert info.converged and info.iterations == ito, (kw, info.iterations, ito)
        assert np.abs(host(x)[:m] - xo).max() <= 1e-14 * np.abs(xo).max(), kw
        if kw.get("build_f32"):
            x.zero_()
            info = H.pcg_reliable(dev(b), x, rel_tol=1e-10, abs_tol=1e-300, max_iter=50)
            assert info.converged, info.iterations
            assert np.abs(host(x)[:m] - xo).max() <= 1e-10 * np.abs(xo).max()
        H.close()


@pytest.mark.parametrize("name,full", [("2d8", True), ("3d8", True), ("diff12", True),
                                       ("2d128", False), ("3d16", False)])
def test_hierarchy_matches_reference_golden(h, golden, name, full):
    g = golden(name)
    rp, ci, v, m = grid(name)
    log(f"hierarchy {name}")
    H = _hier(h, CSR(rp, ci, v, m, m))
    assert H.levels == int(g["levels"])
    for lvl in range(H.levels + 1):
        for kind in "APR":
            key = f"{kind}{lvl}"
            if kind != "A" and lvl == H.levels:
                continue
            grp, gci, gv, gm, gn = H.level(lvl, kind)
            assert list(g[key + "_dims"]) == [gm, gn, len(gci)], key
            if full:
                assert np.array_equal(grp, g[key + "_rp"]), key
                assert np.array_equal(gci, g[key + "_ci"]), key
                assert bits_equal(gv, g[key + "_v"]), key
            else:
                assert list(g[key + "_sha"]) == [sha(grp), sha(gci), sha(gv)], key


def test_hierarchy_matches_oracle_3d48(h, oracle):
    rp, ci, v, m = pb.poisson3d(48)
    A = CSR(rp, ci, v, m, m)
    log("hierarchy 3d48")
    H = _hier(h, A)
    ho = oracle.saamg_setup(A)
    assert H.levels == ho.levels
    for lvl in range(H.levels + 1):
        for kind in "APR":
            if kind != "A" and lvl == H.levels:
                continue
            assert csr_equal(H.level(lvl, kind), ho.get(lvl, kind)), (lvl, kind)


@pytest.mark.parametrize("name", ["2d8", "3d8", "diff12", "2d128", "3d16"])
def test_vcycle_bitwise_reference_golden(h, golden, name):
    """One V-cycle (1+1 Jacobi) with sequential row sums and the reference's
    coarse solve reproduces the reference's preconditioner output bit for bit."""
    g = golden(name)
    rp, ci, v, m = grid(name)
    log(f"vcycle {name}")
    H = _hier(h, CSR(rp, ci, v, m, m), exact=1, coarse_solver=0)
    b = dev(g["b"])
    z = empty(m, torch.float64)
    H.precond(b, z, 1, 1)
    assert bits_equal(host(z), g["precond_z"])


@pytest.mark.parametrize("n1,n2", [(0, 1), (2, 0), (2, 3)])
def test_vcycle_sweep_counts_bitwise(h, oracle, n1, n2):
    rp, ci, v, m = pb.poisson3d(14)
    A = CSR(rp, ci, v, m, m)
    ho = oracle.saamg_setup(A)
    r = pb.uniform(m, 4)
    log(f"vcycle n1={n1} n2={n2}")
    H = _hier(h, A, exact=1, coarse_solver=0)
    R = dev(r)
    z = empty(m, torch.float64)
    H.precond(R, z, n1, n2)
    assert bits_equal(host(z), ho.precond(r, n1, n2, 0))


@pytest.mark.parametrize("name,max_levels", [("3d16", 20), ("diff12", 20), ("2d128", 20),
                                             ("3d8", 3), ("3d8", 2), ("2d128", 3),
                                             ("diff12", 4)])
def test_vcycle_fast_matches_oracle(h, oracle, name, max_levels):
    """One fast-path V-cycle (1 + 1 Jacobi; the coarse levels in the
    single-workgroup LDS tail where they fit; coarsest 1x1 or dense inverse)
    against the oracle's V-cycle with its Cholesky coarsest solve: equal up
    to summation order inside rows."""
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    log(f"vcycle fast {name} max_levels={max_levels}")
    H = _hier(h, A, max_levels=max_levels)
    ho = oracle.saamg_setup(A, max_level=max_levels)
    assert H.levels == ho.levels
    r = np.random.default_rng(6).uniform(-1, 1, m)
    R = dev(r)
    z = empty(m, torch.float64)
    H.precond(R, z, 1, 1)
    zo = ho.precond(r, 1, 1, 1)
    assert np.linalg.norm(host(z) - zo) <= 1e-10 * np.linalg.norm(zo)


@pytest.mark.parametrize("name", ["2d128", "3d16", "diff12"])
def test_pcg_matches_reference(h, golden, name):
    """FP64 SA-AMG-PCG: the GPU reductions sum in another order than the
    host, so the trajectory differs by rounding only: same iteration count
    (+-1), solution within 1e-10 relative of the reference's."""
    g = golden(name)
    rp, ci, v, m = grid(name)
    log(f"pcg {name}")
    H = _hier(h, CSR(rp, ci, v, m, m))
    b = dev(g["b"])
    for tag, tol in (("8", 1e-8), ("10", 1e-10)):
        x = empty(m, torch.float64)
        info = H.pcg(b, x, rel_tol=tol, abs_tol=1e-300, max_iter=1000)
        assert info.converged
        assert abs(info.iterations - int(g[f"pcg{tag}_iters"])) <= 1
        xr = g[f"pcg{tag}_x"]
        xg = host(x)
        assert np.linalg.norm(xg - xr) <= 1e-10 * np.linalg.norm(xr) * (1e2 if tag == "8" else 1)
        A = CSR(rp, ci, v, m, m)
        res = g["b"] - A.to_scipy() @ xg
        assert np.abs(res).max() <= 1.01 * tol * np.abs(g["b"]).max()


def test_pcg_3d64_against_oracle(h, oracle):
    rp, ci, v, m = pb.poisson3d(64)
    A = CSR(rp, ci, v, m, m)
    xs = pb.uniform(m, seed=1)
    b = pb.rhs_from_solution(rp, ci, v, xs)
    log("pcg 3d64")
    ho = oracle.saamg_setup(A)
    xo, ito, _ = ho.pcg(b, rel_tol=1e-10, coarse_solver=1)
    H = _hier(h, A)
    B = dev(b)
    x = empty(m, torch.float64)
    info = H.pcg(B, x, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    xg = host(x)
    assert info.converged and abs(info.iterations - ito) <= 1
    assert np.linalg.norm(xg - xo) <= 1e-10 * np.linalg.norm(xo) * 10
    assert np.linalg.norm(xg - xs) <= 1e-8 * np.linalg.norm(xs)


@pytest.mark.parametrize("name", ["3d16", "diff12", "2d128"])
def test_mixed_precision_ir(h, oracle, name):
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=1))
    log(f"ir {name}")
    H = _hier(h, A, build_f32=1)
    B = dev(b)
    x = empty(m, torch.float64)
    info = H.ir(B, x, rel_tol=1e-10, abs_tol=1e-300, inner_rel_tol=1e-4)
    assert info.converged and info.outer_iterations >= 2
    xg = host(x)
    r = oracle.residual(A, xg, b)
    assert np.abs(r).max() <= 1e-10 * np.abs(b).max()
    xo, _, _ = oracle.saamg_setup(A).pcg(b, rel_tol=1e-12, coarse_solver=1)
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["3d16", "diff12", "2d128"])
def test_mixed_precision_pcg_f32_vcycle(h, oracle, name):
    """FP64 PCG with the V-cycle on FP32 levels: the Krylov recurrence is
    FP64, so the solution reaches the FP64 solver's accuracy and the
    iteration count stays within a few of the FP64 PCG's."""
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=1))
    log(f"pcg_mixed {name}")
    H = _hier(h, A, build_f32=1)
    B = dev(b)
    x = empty(m, torch.float64)
    info = H.pcg_mixed(B, x, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    x64 = empty(m, torch.float64)
    info64 = H.pcg(B, x64, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    assert info.converged and info64.converged
    assert abs(info.iterations - info64.iterations) <= 3, (info.iterations, info64.iterations)
    xg = host(x)
    r = oracle.residual(A, xg, b)
    assert np.abs(r).max() <= 1e-9 * np.abs(b).max()
    xo, _, _ = oracle.saamg_setup(A).pcg(b, rel_tol=1e-12, coarse_solver=1)
    assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo)


@pytest.mark.parametrize("name", ["3d16", "diff12", "2d128"])
def test_mixed_precision_pcg_reliable(h, oracle, name):
    """FP32 PCG and V-cycle with FP64 reliable updates: convergence is declared
    on the true FP64 residual, so it meets the tolerance; the Krylov space is
    kept across updates, so the iteration count stays within a few of the
    FP64 PCG's, and the solution matches the FP64 oracle's to 1e-8."""
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=1))
    log(f"pcg_reliable {name}")
    H = _hier(h, A, build_f32=1)
    B = dev(b)
    x64 = empty(m, torch.float64)
    info64 = H.pcg(B, x64, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
    xo, _, _ = oracle.saamg_setup(A).pcg(b, rel_tol=1e-12, coarse_solver=1)
    for delta in (1e-2, 1e-4):
        x = empty(m, torch.float64)
        info = H.pcg_reliable(B, x, delta=delta, rel_tol=1e-10, abs_tol=1e-300, max_iter=500)
        assert info.converged and info.outer_iterations >= 1, delta
        assert abs(info.iterations - info64.iterations) <= 3, (delta, info.iterations,
                                                              info64.iterations)
        xg = host(x)
        r = oracle.residual(A, xg, b)
        assert np.abs(r).max() <= 1e-10 * np.abs(b).max() * 1.01, delta
        assert np.linalg.norm(xg - xo) <= 1e-8 * np.linalg.norm(xo), delta


def test_pcg_reliable_scale_and_initial_guess(h, oracle):
    """The FP32 residual is the FP64 one scaled by 1/||r0||: right-hand sides
    of 1e-30 and 1e+30 converge like one of order 1, and a nonzero initial
    guess is corrected, not replaced."""
    rp, ci, v, m = grid("3d16")
    A = CSR(rp, ci, v, m, m)
    H = _hier(h, A, build_f32=1)
    b1 = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=3))
    x0 = pb.uniform(m, seed=4)
    its = []
    for scale in (1e-30, 1.0, 1e30):
        b = b1 * scale
        x = dev(x0 * scale)
        info = H.pcg_reliable(dev(b), x, rel_tol=1e-10, abs_tol=0.0, max_iter=500)
        r0 = np.abs(oracle.residual(A, x0 * scale, b)).max()
        r = np.abs(oracle.residual(A, host(x), b)).max()
        assert info.converged and r <= 1e-10 * r0, scale
        its.append(info.iterations)
    assert max(its) - min(its) <= 1, its


def test_mixed_requires_f32_levels(h):
    rp, ci, v, m = grid("2d8")
    H = _hier(h, CSR(rp, ci, v, m, m))
    B = dev(np.ones(m))
    x = empty(m, torch.float64)
    with pytest.raises(lal.LalError):
        H.pcg_mixed(B, x)
    with pytest.raises(lal.LalError):
        H.ir(B, x)
    with pytest.raises(lal.LalError):
        H.pcg_reliable(B, x)


def test_sddmm(h):
    A = random_csr(np.random.default_rng(8), 200, 200, 9)
    d = DevCSR(A)
    u = np.random.default_rng(1).uniform(-1, 1, 200)
    w = np.random.default_rng(2).uniform(-1, 1, 200)
    U, W = dev(u), dev(w)
    out = empty(A.nnz, torch.float64)
    lal.call("lal_sddmm_f64", h.h, A.m, A.nnz, lal.ptr(d.row_ptr), lal.ptr(d.col_ind), -1.0,
             lal.ptr(U), lal.ptr(W), lal.ptr(out))
    rows = np.repeat(np.arange(200), np.diff(A.row_ptr))
    assert bits_equal(host(out)[: A.nnz], (-1.0 * u[rows]) * w[A.col_ind])


def test_sixteen_bit_columns(h, oracle):
    """The fast path reads a hierarchy operator's columns as 16-bit offsets
    from each row block's smallest column when no block spans more than
    65535 columns, as 3D Poisson's finest operator does.  One weak coupling
    100,000 columns away (0 <-> 100,000, -1e-3: symmetric, and far smaller
    than the operator's smallest eigenvalue, 0.0123, so A stays SPD) makes
    two of the finest operator's blocks span too far: it keeps int32
    columns.  Either way PCG takes the oracle's iterations (within one)."""
    for far in (False, True):
        rp, ci, v, m = pb.poisson3d(48)
        if far:
            A = CSR(rp, ci, v, m, m).to_scipy().tolil()
            A[0, 100000] = A[100000, 0] = -1e-3
            A = A.tocsr()
            A.sort_indices()
            rp, ci, v = (A.indptr.astype(np.int32), A.indices.astype(np.int32),
                         A.data.astype(np.float64))
        Ao = CSR(rp, ci, v, m, m)
        H = _hier(h, Ao)
        widths = [H.column_bytes(l, k) for l in range(H.levels + 1) for k in "APR"
                  if k == "A" or l < H.levels]
        log(f"16-bit columns, far coupling {far}: widths {widths}")
        assert widths[0] == (4 if far else 2), widths
        b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=5))
        x = empty(m, torch.float64)
        info = H.pcg(dev(b), x, rel_tol=1e-10, abs_tol=1e-300, max_iter=200)
        _, ito, _ = oracle.saamg_setup(Ao).pcg(b, rel_tol=1e-10)
        assert info.converged and abs(info.iterations - ito) <= 1, (far, info.iterations, ito)
        H.close()


@pytest.mark.parametrize("nx,far,width", [(256, 65535, 2), (257, 65536, 4)])
def test_sixteen_bit_columns_span_limit(h, oracle, nx, far, width):
    """The 16-bit limit itself: 2D Poisson on nx x 256 with one weak coupling
    between rows 0 and `far`.  With 256 x 256 rows and far = 65535 (the last
    row) the two blocks holding those rows span exactly 65535 columns, so the
    finest operator stores offsets up to 65535, the largest 16 bits hold;
    with far = 65536 they span one more and the operator keeps int32
    columns.  The coupling, -1e-4, is below the operator's smallest
    eigenvalue (about 3.0e-4), so A stays SPD.  Either way PCG matches the
    oracle (iterations within one, solution to 1e-8)."""
    rp, ci, v, m = pb.poisson2d(nx, 256)
    A = CSR(rp, ci, v, m, m).to_scipy().tolil()
    A[0, far] = A[far, 0] = -1e-4
    A = A.tocsr()
    A.sort_indices()
    rp, ci, v = A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(np.float64)
    Ao = CSR(rp, ci, v, m, m)
    H = _hier(h, Ao)
    log(f"16-bit span limit: m={m} far={far} width {H.column_bytes(0, 'A')}")
    assert H.column_bytes(0, "A") == width
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=6))
    x = empty(m, torch.float64)
    info = H.pcg(dev(b), x, rel_tol=1e-10, abs_tol=1e-300, max_iter=300)
    xo, ito, _ = oracle.saamg_setup(Ao).pcg(b, rel_tol=1e-10)
    assert info.converged and abs(info.iterations - ito) <= 1, (info.iterations, ito)
    assert np.linalg.norm(host(x) - xo) <= 1e-8 * np.linalg.norm(xo)
    H.close()
