"""Write the golden fixtures in tests/golden/ from LinAlgLib's own host library.

Runs only in a container that holds the reference sources: it needs
oracle/_ref/libref.so, which `make -C oracle ref` compiles from
/root/reference/library/src.  The fixtures are data only (inputs and the
reference's outputs); the tests read them without the reference present.

    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from linalglib_amd import problems as pb  # noqa: E402
from oracle.oracle import CSR, Reference  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def grid(name):
    if name == "2d8":
        return pb.poisson2d(8)
    if name == "3d8":
        return pb.poisson3d(8)
    if name == "2d128":
        return pb.poisson2d(128)
    if name == "3d16":
        return pb.poisson3d(16)
    if name == "diff12":
        return diffusion12()
    raise KeyError(name)


def diffusion12():
    """12^3 diffusion with log-normal cell conductivities (sigma = 1.5), so
    that strength marks a mix of strong and weak couplings."""
    k = np.exp(1.5 * pb.normal(12 ** 3, seed=7))
    return pb.diffusion3d(12, cell_kappa=k)


def hierarchy_case(ref, name, full):
    rp, ci, v, m = grid(name)
    A = CSR(rp, ci, v, m, m)
    h = ref.saamg_setup(A)
    out = {"levels": np.int32(h.levels)}
    for lvl in range(h.levels + 1):
        for kind in "APR":
            M = h.get(lvl, kind)
            if M is None:
                continue
            key = f"{kind}{lvl}"
            out[key + "_dims"] = np.array([M.m, M.n, M.nnz], dtype=np.int64)
            if full:
                out[key + "_rp"] = M.row_ptr
                out[key + "_ci"] = M.col_ind
                out[key + "_v"] = M.val
            else:
                out[key + "_sha"] = np.array(
                    [sha(M.row_ptr), sha(M.col_ind), sha(M.val)], dtype="U64")
    xs = pb.uniform(m, seed=1)
    b = pb.rhs_from_solution(rp, ci, v, xs)
    out["b"] = b
    z = h.precond(b, 1, 1)
    out["precond_z"] = z
    for tol, tag in ((1e-8, "8"), (1e-10, "10")):
        x, it = h.pcg(b, rel_tol=tol, abs_tol=1e-300, max_iter=1000, n1=1, n2=1)
        out[f"pcg{tag}_iters"] = np.int32(it)
        out[f"pcg{tag}_x"] = x
    np.savez_compressed(os.path.join(OUT, f"golden_{name}.npz"), **out)
    print(name, "levels", h.levels, "iters", int(out["pcg8_iters"]), int(out["pcg10_iters"]))


def random_csr(rng, m, n, per_row, vals="uniform"):
    rows = []
    for i in range(m):
        k = int(rng.integers(0, per_row + 1))
        rows.append(np.sort(rng.choice(n, size=min(k, n), replace=False)))
    rp = np.zeros(m + 1, dtype=np.int32)
    rp[1:] = np.cumsum([len(r) for r in rows])
    ci = np.concatenate(rows).astype(np.int32) if rp[-1] else np.zeros(0, np.int32)
    if vals == "uniform":
        v = rng.uniform(-1, 1, rp[-1])
    else:  # small integers: products cancel to exact zeros that must be kept
        v = rng.integers(-2, 3, rp[-1]).astype(np.float64)
    return CSR(rp, ci, v, m, n)


def ops_case(ref):
    rng = np.random.default_rng(1234567)
    out = {}
    # SpMV with alpha/beta, including empty rows
    A = random_csr(rng, 300, 200, 12)
    x = rng.uniform(-1, 1, 200)
    y0 = rng.uniform(-1, 1, 300)
    out.update(mv_rp=A.row_ptr, mv_ci=A.col_ind, mv_v=A.val, mv_x=x, mv_y0=y0)
    for tag, a, bb in (("a1b0", 1.0, 0.0), ("a1b1", 1.0, 1.0), ("am2b05", -2.0, 0.5)):
        out["mv_" + tag] = ref.csrmv(A, x, a, bb, y0)
    # residual and dot
    S = random_csr(rng, 250, 250, 9)
    xs = rng.uniform(-1, 1, 250)
    bs = rng.uniform(-1, 1, 250)
    out.update(res_rp=S.row_ptr, res_ci=S.col_ind, res_v=S.val, res_x=xs, res_b=bs)
    out["res_r"] = ref.residual(S, xs, bs)
    dx = rng.uniform(-1, 1, 5000)
    dy = rng.uniform(-1, 1, 5000)
    ref.set_threads(1)
    out.update(dot_x=dx, dot_y=dy, dot_1t=np.float64(ref.dot(dx, dy)))
    # SpGEMM (values with cancellations) and transpose
    G1 = random_csr(rng, 120, 90, 8, vals="int")
    G2 = random_csr(rng, 90, 110, 8, vals="int")
    G = ref.csrgemm(G1, G2)
    out.update(g1_rp=G1.row_ptr, g1_ci=G1.col_ind, g1_v=G1.val,
               g2_rp=G2.row_ptr, g2_ci=G2.col_ind, g2_v=G2.val,
               g_rp=G.row_ptr, g_ci=G.col_ind, g_v=G.val)
    H1 = random_csr(rng, 80, 70, 10)
    H2 = random_csr(rng, 70, 60, 10)
    H = ref.csrgemm(H1, H2)
    out.update(h1_rp=H1.row_ptr, h1_ci=H1.col_ind, h1_v=H1.val,
               h2_rp=H2.row_ptr, h2_ci=H2.col_ind, h2_v=H2.val,
               h_rp=H.row_ptr, h_ci=H.col_ind, h_v=H.val)
    T = ref.transpose(H1)
    out.update(t_rp=T.row_ptr, t_ci=T.col_ind, t_v=T.val)
    # strength and PMIS aggregates on a heterogeneous operator
    rp, ci, v, m = diffusion12()
    D = CSR(rp, ci, v, m, m)
    for eps, tag in ((0.08, "08"), (0.25, "25")):
        conn = ref.strength(D, eps)
        out["str_" + tag] = conn
        out["agg_" + tag] = ref.aggregates(D, conn)
    np.savez_compressed(os.path.join(OUT, "golden_ops.npz"), **out)
    print("ops written")


def cycles_case(ref):
    """amg_solve with every cycle (V, W, F) and smoother (Jacobi, GS, SGS, SOR,
    SSOR) on the 2D 8x8 and 3D 8^3 Poisson grids: cycle counts and solutions."""
    from oracle.oracle import CSR

    out = {}
    for name in ("2d8", "3d8"):
        rp, ci, v, m = grid(name)
        b = np.random.default_rng(1).uniform(-1, 1, m)
        h = ref.saamg_setup(CSR(rp, ci, v, m, m))
        out[f"{name}_b"] = b
        for cyc in range(3):
            for sm in range(5):
                x, k = h.amg_solve(b, n1=1, n2=1, cycle=cyc, smoother=sm, rel_tol=1e-9,
                                   max_cycle=40)
                out[f"{name}_c{cyc}_s{sm}_x"] = x
                out[f"{name}_c{cyc}_s{sm}_k"] = np.array([k])
    np.savez_compressed(os.path.join(OUT, "golden_cycles.npz"), **out)
    print("cycles written")


def ua_case(ref):
    """uaamg_setup hierarchies: full arrays for the 2D 8x8 and 3D 8^3 grids,
    SHA-256 of every level for 2D 128^2 and 3D 16^3."""
    from oracle.oracle import CSR

    out = {}
    for name, full in (("2d8", True), ("3d8", True), ("2d128", False), ("3d16", False)):
        rp, ci, v, m = grid(name)
        h = ref.uaamg_setup(CSR(rp, ci, v, m, m))
        out[f"{name}_levels"] = np.array([h.levels])
        for lvl in range(h.levels + 1):
            for kind in ("A", "P", "R") if lvl < h.levels else ("A",):
                M = h.get(lvl, kind)
                key = f"{name}_{kind}{lvl}"
                out[key + "_dims"] = np.array([M.m, M.n, M.nnz])
                if full:
                    out[key + "_rp"], out[key + "_ci"], out[key + "_v"] = M.row_ptr, M.col_ind, M.val
                else:
                    out[key + "_sha"] = np.frombuffer(
                        (sha(M.row_ptr) + sha(M.col_ind) + sha(M.val)).encode(), dtype=np.uint8)
    np.savez_compressed(os.path.join(OUT, "golden_ua.npz"), **out)
    print("ua written")


def krylov_case(ref):
    """bicgstab_solver with no, Jacobi and SA-AMG preconditioning on two grids:
    iteration counts and solutions (the reference's GMRES does not run: its
    build() leaves the residual vectors unsized)."""
    from oracle.oracle import CSR

    out = {}
    for name in ("2d8", "3d8"):
        rp, ci, v, m = grid(name)
        A = CSR(rp, ci, v, m, m)
        b = np.random.default_rng(1).uniform(-1, 1, m)
        h = ref.saamg_setup(A)
        out[f"{name}_b"] = b
        for pc in range(3):
            x, it = ref.bicgstab(A, b, precond=pc, hier=h, rel_tol=1e-10, max_iter=500)
            out[f"{name}_p{pc}_x"] = x
            out[f"{name}_p{pc}_it"] = np.array([it])
    np.savez_compressed(os.path.join(OUT, "golden_krylov.npz"), **out)
    print("krylov written")


def op_precond_matrix():
    """8^3 diffusion with log-normal cell conductivities (sigma = 1), the
    matrix a preconditioner is built for while the Krylov operator is 3D 8^3
    Poisson."""
    k = np.exp(1.0 * pb.normal(8 ** 3, seed=5))
    return pb.diffusion3d(8, cell_kappa=k)


def cg_case(ref):
    """cg_solver::solve without a preconditioner (solve_nonprecond, whose
    restart block runs on every iteration) on 3D 6^3 and 2D 12^2 Poisson at
    two tolerances: iteration counts and solutions."""
    out = {}
    for name, gen in (("3d6", lambda: pb.poisson3d(6)), ("2d12", lambda: pb.poisson2d(12))):
        rp, ci, v, m = gen()
        A = CSR(rp, ci, v, m, m)
        b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=3))
        out[f"{name}_b"] = b
        for tag, rt, at in (("6", 1e-6, 1e-6), ("10", 1e-10, 1e-300)):
            x, it = ref.cg(A, b, rel_tol=rt, abs_tol=at, max_iter=2000)
            out[f"{name}_t{tag}_x"] = x
            out[f"{name}_t{tag}_it"] = np.array([it])
    # the C++ API program's case: b = 1 on 3D 6^3 Poisson, rel_tol 1e-6
    rp, ci, v, m = pb.poisson3d(6)
    x, it = ref.cg(CSR(rp, ci, v, m, m), np.ones(m), rel_tol=1e-6, abs_tol=1e-300,
                   max_iter=5000)
    out.update(ones6_x=x, ones6_it=np.array([it]))
    # cg_solver::solve(A, x, b, precond) with the SA-AMG V-cycle of a hierarchy
    # built for another operator of the same size (log-normal diffusion 8^3)
    # and the Krylov products on 3D 8^3 Poisson
    rp, ci, v, m = pb.poisson3d(8)
    drp, dci, dv, _ = op_precond_matrix()
    h = ref.saamg_setup(CSR(drp, dci, dv, m, m))
    b = pb.rhs_from_solution(rp, ci, v, pb.uniform(m, seed=2))
    x, it = h.pcg_op(CSR(rp, ci, v, m, m), b, rel_tol=1e-10)
    out.update(op_b=b, op_x=x, op_it=np.array([it]))
    np.savez_compressed(os.path.join(OUT, "golden_cg.npz"), **out)
    print("cg written")


def main():
    ref = Reference()
    ref.set_threads(1)
    if len(sys.argv) > 1 and sys.argv[1] == "cg":
        cg_case(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "krylov":
        krylov_case(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "ua":
        ua_case(ref)
        return
    if len(sys.argv) > 1 and sys.argv[1] == "cycles":
        cycles_case(ref)
        return
    hierarchy_case(ref, "2d8", full=True)
    hierarchy_case(ref, "3d8", full=True)
    hierarchy_case(ref, "diff12", full=True)
    hierarchy_case(ref, "2d128", full=False)
    hierarchy_case(ref, "3d16", full=False)
    ops_case(ref)
    cycles_case(ref)
    ua_case(ref)
    krylov_case(ref)
    cg_case(ref)


if __name__ == "__main__":
    main()
