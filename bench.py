#!/usr/bin/env python3
"""SA-AMG-PCG on one MI355X: solve seconds to rel. residual 1e-10, mixed
FP32/FP64 against plain FP64, and standalone CSR SpMV GB/s.

Workload: 3D 7-point Poisson on a 128^3 grid (n = 2,097,152, nnz = 14,581,760),
FP64 CSR with int32 indices, x* ~ U[-1,1] (seed 1), b = A x*, x0 = 0.  SA-AMG
hierarchy (strength eps = 0.08 halved per level, smoothed prolongator with
relax = 2/3), one V-cycle with 1 pre + 1 post undamped Jacobi sweep per
preconditioner application, PCG to ||r||_inf / ||r0||_inf <= 1e-10.

A step is one complete solve from x = 0 with the hierarchy resident in HBM
(setup is timed once and reported separately).  `value` is the mixed
precision solve time of the scheme --mixed names; three mixed schemes are
timed:
  reliable  FP32 PCG and V-cycle (FP64 inner-product accumulation) with FP64
            reliable updates of x and the true residual every 1e-2 of
            residual reduction (the default, `value`)
  vcycle    FP64 PCG (x, r, p, A p, inner products) with the V-cycle applied
            on FP32 copies of the hierarchy
  ir        FP32 SA-AMG-PCG to 1e-4 inside FP64 iterative refinement to 1e-10
and the plain FP64 solve time is reported beside them.

The CPU baseline runs LinAlgLib's own host library on the same problem with
every host core the job was given (OMP_NUM_THREADS, else the affinity mask)
and again with one thread.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nx", type=int, default=128)
    ap.add_argument("--spmv-calls", type=int, default=200)
    ap.add_argument("--spmv-only", action="store_true",
                    help="only the standalone SpMV loop (for counter runs)")
    ap.add_argument("--mixed", choices=("vcycle", "ir", "reliable"), default="reliable",
                    help="mixed scheme reported as value")
    ap.add_argument("--config", choices=("c2", "c1", "c3", "c4", "c5", "amg", "krylov"),
                    default="c2",
                    help="c2 (default): 3D Poisson 128^3, the headline line; the others "
                         "print one JSON line each for the remaining workloads: c1 2D "
                         "Poisson 128^2 to 1e-8, c3 log-normal FV diffusion 160^3, c4 96^3 "
                         "face-conductance gradient, c5 kNN graph Laplacian (2M points), "
                         "amg: AMG as a solver (V/W/F cycles, Jacobi/GS/SSOR) on 128^3, "
                         "krylov: BiCGSTAB / GMRES(30) with Jacobi and SA-AMG "
                         "preconditioning on 128^3 Poisson and 1024^2 convection-diffusion")
    ap.add_argument("--hbm-ceiling", action="store_true",
                    help="only time a 2 GiB streaming reduction and a 1 GiB copy (torch's "
                         "kernels): the HBM bandwidth the SpMV fraction is set against")
    ap.add_argument("--setup-only", action="store_true",
                    help="only build the SA-AMG hierarchy of the --nx Poisson problem "
                         "--steps times (after --warmup builds) and print the setup times")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    return ap.parse_args()


def spmv_bytes(m, n, nnz, vbytes, xbytes, beta=False, cbytes=4):
    """Algorithmic bytes of y = A x: values + columns (cbytes each: int32, or
    the 16-bit offsets a hierarchy's operators may carry) per nonzero, int32
    row pointers, x read once, y written (and read when beta != 0)."""
    return nnz * (vbytes + cbytes) + 4 * (m + 1) + xbytes * n + xbytes * m * (2 if beta else 1)


def solve_bytes(levels, iters, vbytes):
    """Algorithmic bytes of a SA-AMG-PCG solve: per iteration one SpMV with A,
    one V-cycle, two dot products, one inf-norm and three axpy-class updates
    (vectors of vbytes each); levels = [(m, nnz_A, nnz_P, cA, cP, cR)] finest
    first (cX: bytes per column index the kernels read for A, P and R =
    P^T), the last entry being the coarsest operator."""
    v = vbytes
    m0, nnz0, _, c0, _, _ = levels[0]
    prec = 0
    for (m, nnz, nnz_p, ca, cp, cr), nxt in zip(levels[:-1], levels[1:]):
        mc = nxt[0]
        jac0 = 3 * v * m                                   # x = b / a_ii
        res = nnz * (v + ca) + 4 * (m + 1) + v * m + 2 * v * m
        restr = spmv_bytes(mc, m, nnz_p, v, v, cbytes=cr)
        prol = spmv_bytes(m, mc, nnz_p, v, v, beta=True, cbytes=cp)
        post = nnz * (v + ca) + 4 * (m + 1) + 3 * v * m
        prec += jac0 + res + restr + prol + post
    per_iter = (spmv_bytes(m0, m0, nnz0, v, v, cbytes=c0) + prec + 2 * 2 * v * m0 + v * m0 +
                3 * 3 * v * m0)
    return iters * per_iter


def reliable_update_bytes(m, nnz, cbytes=4):
    """Algorithmic bytes of one FP64 update of the reliable-update scheme:
    x += ||r0|| y (y FP32 read, x FP64 read and written), y = 0, r = b - A x in
    FP64 (cbytes per column index), its inf-norm, and the scaled FP32 copy of r."""
    return 20 * m + 4 * m + (nnz * (8 + cbytes) + 4 * (m + 1) + 24 * m) + 8 * m + 12 * m


def measured_traffic():
    """HBM bytes per FP64 SpMV launch at 128^3 from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE passes (profiles/spmv_poisson128_hbm_traffic.md)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "spmv_poisson128_traffic.txt")
    try:
        kv = dict(line.strip().split("=", 1) for line in open(path) if "=" in line)
        return int(float(kv["traffic_bytes_fp64"]))
    except (OSError, KeyError, ValueError):
        return None


def host_threads():
    """Host cores this job may use: OMP_NUM_THREADS when the launcher set it
    (the share of the machine given to one GPU), else the affinity mask."""
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return int(env)
    return len(os.sched_getaffinity(0))


def cpu_baseline(rp, ci, v, m, b, threads, tol=1e-10, name=None, one_thread=True):
    """LinAlgLib's own host path (oracle/_ref, compiled from its sources) on
    the same problem when it was built, else the C oracle; SA setup and one
    SA-AMG-PCG solve to `tol`, timed on this machine's host cores (`threads`,
    then one thread on the same hierarchy when one_thread is set)."""
    here = os.path.dirname(os.path.abspath(__file__))
    sys.path.insert(0, here)
    from oracle.oracle import CSR, Oracle, Reference

    A = CSR(rp, ci, v, m, m)
    if Reference.available():
        lib, kind = Reference(), "reference"
    else:
        lib, kind = Oracle(), "port"
    lib.set_threads(threads)
    t0 = time.perf_counter()
    h = lib.saamg_setup(A)
    t1 = time.perf_counter()
    out = h.pcg(b, rel_tol=tol, abs_tol=1e-300, max_iter=1000)
    t2 = time.perf_counter()
    iters = out[1]
    solve1 = None
    if one_thread and threads > 1:
        lib.set_threads(1)
        t3 = time.perf_counter()
        h.pcg(b, rel_tol=tol, abs_tol=1e-300, max_iter=1000)
        solve1 = round(time.perf_counter() - t3, 4)
        lib.set_threads(threads)
    return {
        "value": round(t2 - t1, 4),
        "unit": f"s per SA-AMG-PCG solve to {tol:g} (FP64)",
        "cores": threads,
        "kind": kind,
        "sample": (f"one full solve of the same "
                   f"{name or f'3D Poisson {round(m ** (1 / 3))}^3'} problem "
                   f"(setup {t1 - t0:.2f} s excluded, {iters} PCG iterations, "
                   f"{h.levels} levels), OMP_NUM_THREADS={threads}"),
        "setup_s": round(t1 - t0, 3),
        "iterations": iters,
        "value_1_thread": solve1 if solve1 is not None else round(t2 - t1, 4),
    }


def timed_solves(H, B, X, kinds, steps, warmup, rel_tol):
    """Median seconds and info of each solve kind ("fp64", "vcycle", "ir")
    from x = 0, with the hierarchy resident."""
    import torch

    def run(kind):
        X.zero_()
        if kind == "fp64":
            return H.pcg(B, X, rel_tol=rel_tol, abs_tol=1e-300, max_iter=1000)
        if kind == "vcycle":
            return H.pcg_mixed(B, X, rel_tol=rel_tol, abs_tol=1e-300, max_iter=1000)
        if kind == "reliable":
            return H.pcg_reliable(B, X, delta=1e-2, rel_tol=rel_tol, abs_tol=1e-300,
                                  max_iter=1000)
        return H.ir(B, X, rel_tol=rel_tol, abs_tol=1e-300, inner_rel_tol=1e-4,
                    inner_max_iter=200, max_outer=30)

    out = {}
    for kind in kinds:
        for _ in range(warmup):
            run(kind)
        torch.cuda.synchronize()
        times = []
        for _ in range(steps):
            s0 = time.perf_counter()
            info = run(kind)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - s0)
        out[kind] = {"median_s": statistics.median(times), "iters": info.iterations,
                     "outer": info.outer_iterations, "converged": bool(info.converged),
                     "x": X.clone()}
    return out


def run_config(args):
    """The workloads other than the headline 128^3 Poisson: one JSON line each
    with FP64 and mixed solve times (and, for c4, the gradient)."""
    import torch

    from linalglib_amd import lal, problems
    from linalglib_amd.torch_solve import saamg_solve

    dev = torch.device("cuda:0")
    h = lal.Handle()
    cfg = args.config
    rng = np.random.default_rng(2024)
    if cfg == "c1":
        rp, ci, v, m = problems.poisson2d(128)
        b = problems.rhs_from_solution(rp, ci, v, problems.uniform(m, seed=1))
        tol, name = 1e-8, "2D 5-point Poisson 128^2 to 1e-8"
    elif cfg == "c3":
        nx = 160
        kappa = np.exp(1.5 * rng.standard_normal((nx, nx, nx)))
        rp, ci, v, m = problems.diffusion3d(nx, cell_kappa=kappa)
        b = rng.standard_normal(m)
        tol, name = 1e-10, "3D FV diffusion 160^3, kappa = exp(N(0, 1.5^2)), to 1e-10"
    elif cfg == "c5":
        n = 2_000_000
        rp, ci, m, E, entry, edge, sign, dpos = problems.knn_laplacian(n, seed=5)
        w = rng.uniform(0.5, 2.0, len(E))
        v = problems.assemble_knn(len(ci), entry, edge, sign, dpos, w)
        b = rng.standard_normal(m)
        tol, name = 1e-10, "kNN graph Laplacian + 1e-2 I, 2M points in [0,1]^3, 12-NN, to 1e-10"
    elif cfg == "amg":
        return run_amg(args)
    elif cfg == "krylov":
        return run_krylov(args)
    else:
        return run_gradient(args)
    nnz = int(rp[-1])
    RP, CI, V, B = (torch.from_numpy(a).to(dev) for a in (rp, ci, v, b))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H = lal.Hierarchy(h, m, nnz, RP, CI, V, build_f32=1)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    X = torch.zeros(m, dtype=torch.float64, device=dev)
    res = timed_solves(H, B, X, ("fp64", "vcycle", "ir", "reliable"), args.steps, args.warmup,
                       tol)
    line = {"config": {"workload": name, "n": m, "nnz": nnz, "levels": H.levels},
            "setup_s": round(setup_s, 4)}
    for kind, r in res.items():
        x = r.pop("x").cpu().numpy()
        rr = b - problems.rhs_from_solution(rp, ci, v, x)
        line[kind] = {"solve_s": round(r["median_s"], 6), "iterations": r["iters"],
                      "outer": r["outer"], "converged": r["converged"],
                      "rel_residual_inf": float(np.abs(rr).max() / np.abs(b).max())}
    line["speedup_vcycle_vs_fp64"] = round(res["fp64"]["median_s"] / res["vcycle"]["median_s"], 3)
    line["speedup_ir_vs_fp64"] = round(res["fp64"]["median_s"] / res["ir"]["median_s"], 3)
    line["speedup_reliable_vs_fp64"] = round(res["fp64"]["median_s"] / res["reliable"]["median_s"],
                                             3)
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        try:
            line["cpu_baseline"] = cpu_baseline(rp, ci, v, m, b, threads, tol, name,
                                                one_thread=False)
        except Exception as e:  # the baseline is context only
            line["cpu_baseline"] = {"value": None, "cores": threads, "sample": f"not run: {e}"}
    print(json.dumps(line), flush=True)


def run_amg(args):
    """AMG as a stand-alone solver on 3D Poisson 128^3: seconds per cycle for
    V/W/F cycles with Jacobi, Gauss-Seidel and SSOR smoothing (fast kernels;
    Gauss-Seidel-type sweeps level-scheduled), cycles to 1e-8, and LinAlgLib's
    host amg_solve per cycle on the same problem (one cycle, bounded sample)."""
    import torch

    from linalglib_amd import lal, problems

    dev = torch.device("cuda:0")
    h = lal.Handle()
    rp, ci, v, m = problems.poisson3d(128)
    b = problems.rhs_from_solution(rp, ci, v, problems.uniform(m, seed=1))
    RP, CI, V, B = (torch.from_numpy(a).to(dev) for a in (rp, ci, v, b))
    H = lal.Hierarchy(h, m, int(rp[-1]), RP, CI, V)
    X = torch.zeros(m, dtype=torch.float64, device=dev)
    line = {"config": {"workload": "AMG as a solver, 3D Poisson 128^3, 1 pre + 1 post sweep",
                       "n": m, "nnz": int(rp[-1]), "levels": H.levels}, "gpu": {}}
    for cyc in ("v", "w", "f"):
        for sm in ("jacobi", "gs", "ssor"):
            # first call builds the Gauss-Seidel level sets; then one timed cycle
            X.zero_()
            H.amg_solve(B, X, 1, 1, cyc, sm, rel_tol=1e-300, abs_tol=1e-300, max_iter=0)
            torch.cuda.synchronize()
            X.zero_()
            t0 = time.perf_counter()
            H.amg_solve(B, X, 1, 1, cyc, sm, rel_tol=1e-300, abs_tol=1e-300, max_iter=0)
            torch.cuda.synchronize()
            rec = {"s_per_cycle": round(time.perf_counter() - t0, 6)}
            if cyc == "v":
                X.zero_()
                t0 = time.perf_counter()
                info = H.amg_solve(B, X, 1, 1, cyc, sm, rel_tol=1e-8, abs_tol=1e-300,
                                   max_iter=60)
                torch.cuda.synchronize()
                rec.update({"cycles_to_1e-8": info.iterations, "converged": bool(info.converged),
                            "solve_s": round(time.perf_counter() - t0, 5)})
            line["gpu"][f"{cyc}_{sm}"] = rec
            print(f"# amg {cyc} {sm}: {rec}", file=sys.stderr, flush=True)
    # unsmoothed aggregation: setup and PCG with its V-cycle
    del H
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    HU = lal.Hierarchy(h, m, int(rp[-1]), RP, CI, V, unsmoothed=True)
    torch.cuda.synchronize()
    ua_setup = time.perf_counter() - t0
    X.zero_()
    HU.pcg(B, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=1000)
    torch.cuda.synchronize()
    X.zero_()
    t0 = time.perf_counter()
    info = HU.pcg(B, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=1000)
    torch.cuda.synchronize()
    line["gpu"]["uaamg"] = {"setup_s": round(ua_setup, 4), "levels": HU.levels,
                            "pcg_solve_s": round(time.perf_counter() - t0, 5),
                            "pcg_iterations": info.iterations, "converged": bool(info.converged)}
    print(f"# amg uaamg: {line['gpu']['uaamg']}", file=sys.stderr, flush=True)
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle.oracle import CSR, Oracle, Reference

        threads = args.cpu_threads or host_threads()
        lib = Reference() if Reference.available() else Oracle()
        lib.set_threads(threads)
        print("# amg: host setup and one cycle of each kind on the CPU", file=sys.stderr,
              flush=True)
        t0 = time.perf_counter()
        lib.uaamg_setup(CSR(rp, ci, v, m, m))
        ua_cpu_setup = time.perf_counter() - t0
        hc = lib.saamg_setup(CSR(rp, ci, v, m, m))
        cpu = {}
        for cyc, sm in ((0, 0), (0, 1)):
            t0 = time.perf_counter()
            hc.amg_solve(b, cycle=cyc, smoother=sm, rel_tol=1e-300, max_cycle=0)
            cpu[f"{'vwf'[cyc]}_{['jacobi', 'gs'][sm]}"] = round(time.perf_counter() - t0, 4)
        line["cpu_baseline"] = {"s_per_cycle": cpu, "uaamg_setup_s": round(ua_cpu_setup, 3),
                                "cores": threads,
                                "kind": "reference" if Reference.available() else "port",
                                "sample": "one cycle of amg_solve on the same problem"}
    print(json.dumps(line), flush=True)


def run_krylov(args):
    """BiCGSTAB and restarted GMRES(30) beside SA-AMG-PCG: seconds to 1e-10
    and per iteration on 3D Poisson 128^3 (SPD) and on 2D upwind
    convection-diffusion 1024^2 (nonsymmetric), with Jacobi and SA-AMG
    preconditioning; the host reference's BiCGSTAB (and the oracle's GMRES:
    the reference's cannot run) per iteration on the 128^3 problem."""
    import torch

    from linalglib_amd import lal, problems

    dev = torch.device("cuda:0")
    h = lal.Handle()
    line = {"config": {"workload": "Krylov solvers to 1e-10: 3D Poisson 128^3 and 2D "
                                   "convection-diffusion 1024^2 (wind (20, -10), upwind)"},
            "gpu": {}}

    def timed(tag, fn, X):
        X.zero_()
        fn()
        torch.cuda.synchronize()
        X.zero_()
        t0 = time.perf_counter()
        info = fn()
        torch.cuda.synchronize()
        sec = time.perf_counter() - t0
        rec = {"solve_s": round(sec, 5), "iterations": info.iterations,
               "converged": bool(info.converged)}
        if "gmres" in tag:
            rec["arnoldi_steps"] = info.outer_iterations
            rec["s_per_arnoldi_step"] = round(sec / max(info.outer_iterations, 1), 7)
        else:
            rec["s_per_iteration"] = round(sec / (info.iterations + 1), 7)
        line["gpu"][tag] = rec
        print(f"# krylov {tag}: {rec}", file=sys.stderr, flush=True)

    for prob in ("poisson128", "cd1024"):
        if prob == "poisson128":
            rp, ci, v, m = problems.poisson3d(128)
        else:
            rp, ci, v, m = problems.convection_diffusion2d(1024)
        b = problems.rhs_from_solution(rp, ci, v, problems.uniform(m, seed=1))
        nnz = int(rp[-1])
        RP, CI, V, B = (torch.from_numpy(a).to(dev) for a in (rp, ci, v, b))
        D = lal.CsrDescr(h, m, m, nnz, RP, CI)
        A = (RP, CI, V, m, nnz, D)
        H = lal.Hierarchy(h, m, nnz, RP, CI, V)
        X = torch.zeros(m, dtype=torch.float64, device=dev)
        kw = dict(rel_tol=1e-10, abs_tol=1e-300)
        line["config"][prob] = {"n": m, "nnz": nnz, "levels": H.levels}
        if prob == "poisson128":
            timed(f"{prob}_pcg_saamg", lambda: H.pcg(B, X, max_iter=1000, **kw), X)
        timed(f"{prob}_bicgstab_saamg",
              lambda: lal.bicgstab(h, A, B, X, "saamg", H, max_iter=1000, **kw), X)
        timed(f"{prob}_gmres30_saamg",
              lambda: lal.gmres(h, A, B, X, 30, "saamg", H, max_iter=200, **kw), X)
        timed(f"{prob}_bicgstab_jacobi",
              lambda: lal.bicgstab(h, A, B, X, "jacobi", None, max_iter=5000, **kw), X)
        del H, D
        torch.cuda.synchronize()
    if not args.no_cpu_baseline:
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        from oracle.oracle import CSR, Oracle, Reference

        threads = args.cpu_threads or host_threads()
        rp, ci, v, m = problems.poisson3d(128)
        b = problems.rhs_from_solution(rp, ci, v, problems.uniform(m, seed=1))
        Ac = CSR(rp, ci, v, m, m)
        lib = Reference() if Reference.available() else Oracle()
        lib.set_threads(threads)
        print("# krylov: host BiCGSTAB (3 iterations) and GMRES (one 30-step cycle)",
              file=sys.stderr, flush=True)
        hc = lib.saamg_setup(Ac)
        t0 = time.perf_counter()
        lib.bicgstab(Ac, b, precond=2, hier=hc, rel_tol=1e-300, max_iter=2)
        bic = (time.perf_counter() - t0) / 3
        orc = Oracle()
        orc.set_threads(threads)
        ho = orc.saamg_setup(Ac)
        t0 = time.perf_counter()
        orc.gmres(Ac, b, restart=30, precond=2, hier=ho, rel_tol=1e-300, max_iter=0)
        gm = (time.perf_counter() - t0) / 30
        line["cpu_baseline"] = {
            "bicgstab_saamg_s_per_iteration": round(bic, 4), "bicgstab_cores": threads,
            "bicgstab_kind": "reference" if Reference.available() else "port",
            "gmres_saamg_s_per_arnoldi_step": round(gm, 4), "gmres_cores": threads,
            "gmres_kind": "port",
            "sample": "poisson128: 3 BiCGSTAB iterations; one GMRES(30) restart cycle"}
    print(json.dumps(line), flush=True)


def run_gradient(args):
    """96^3 face-conductance diffusion: L = 0.5 ||x - x_t||^2, x = A(k)^-1 b;
    forward and adjoint solves (FP64 and mixed) plus dL/dk, and a central
    difference check of 64 seeded faces with FP64 solves (h = 1e-6) that share
    the hierarchy of the unperturbed operator."""
    import torch

    from linalglib_amd import problems
    from linalglib_amd.torch_solve import saamg_solve

    dev = torch.device("cuda:0")
    nx = 96
    rp, ci, m, nf, entry, face, sign = problems.diffusion_faces(nx)
    rng = np.random.default_rng(96)
    k = np.exp(rng.standard_normal(nf))
    b = rng.standard_normal(m)
    xt = rng.standard_normal(m)
    RP, CI = torch.from_numpy(rp).to(dev), torch.from_numpy(ci).to(dev)
    E, F, S = (torch.from_numpy(a).to(dev) for a in (entry, face, sign))
    Bt, XT = torch.from_numpy(b).to(dev), torch.from_numpy(xt).to(dev)

    def loss_grad(kk, mixed, hierarchy=None):
        K = torch.tensor(kk, device=dev, requires_grad=True)
        val = torch.zeros(len(ci), dtype=torch.float64, device=dev).index_add(0, E, S * K[F])
        x = saamg_solve(val, Bt, RP, CI, mixed=mixed, tol=1e-10, hierarchy=hierarchy)
        L = 0.5 * ((x - XT) ** 2).sum()
        L.backward()
        torch.cuda.synchronize()
        return L.item(), K.grad

    out = {"config": {"workload": "96^3 FV diffusion, k_f = exp(N(0,1)) per face, "
                                  "L = 0.5 ||x - x_t||^2, dL/dk by adjoint + SDDMM",
                      "n": m, "nnz": int(rp[-1]), "faces": nf}}
    grads = {}
    for mixed in (False, "vcycle"):
        loss_grad(k, mixed)
        times = []
        for _ in range(max(1, args.steps)):
            t0 = time.perf_counter()
            L, g = loss_grad(k, mixed)
            times.append(time.perf_counter() - t0)
        grads[mixed] = g.cpu().numpy()
        out["fp64" if mixed is False else "mixed"] = {
            "forward_backward_s": round(statistics.median(times), 5), "loss": L}
    g64, gmx = grads[False], grads["vcycle"]
    out["mixed_grad_rel_diff"] = float(np.linalg.norm(gmx - g64) / np.linalg.norm(g64))
    # an optimisation loop's later steps: the hierarchy built once (here for
    # A(k) itself) preconditions FP64 PCG on the current values, forward and
    # adjoint, so no setup is paid per step
    from linalglib_amd import lal

    hr = lal.Handle()
    V0r = torch.zeros(len(ci), dtype=torch.float64, device=dev).index_add(
        0, E, S * torch.from_numpy(k).to(dev)[F])
    Hr = lal.Hierarchy(hr, m, len(ci), RP, CI, V0r)
    loss_grad(k, False, hierarchy=Hr)
    times = []
    for _ in range(max(1, args.steps)):
        t0 = time.perf_counter()
        L, g = loss_grad(k, False, hierarchy=Hr)
        times.append(time.perf_counter() - t0)
    out["fp64_reused_hierarchy"] = {
        "forward_backward_s": round(statistics.median(times), 5), "loss": L,
        "grad_rel_diff_vs_rebuilt": float(np.linalg.norm(g.cpu().numpy() - g64) /
                                          np.linalg.norm(g64))}
    Hr.close()
    # central differences on 64 seeded faces: two FP64 solves to 1e-13 of
    # A(k +- h e_f) per face, both preconditioned by the hierarchy of A(k), and
    # L(k+h) - L(k-h) = 0.5 (x+ - x-) . (x+ + x- - 2 x_t)
    faces = rng.choice(nf, size=64, replace=False)
    hstep = 1e-6
    hh = lal.Handle()
    V0 = torch.zeros(len(ci), dtype=torch.float64, device=dev).index_add(
        0, E, S * torch.from_numpy(k).to(dev)[F])
    H0 = lal.Hierarchy(hh, m, len(ci), RP, CI, V0)

    def fd_solve(kk):
        Kd = torch.from_numpy(kk).to(dev)
        Vd = torch.zeros(len(ci), dtype=torch.float64, device=dev).index_add(0, E, S * Kd[F])
        xs = torch.zeros(m, dtype=torch.float64, device=dev)
        H0.pcg_op((RP, CI, Vd, m, len(ci), None), Bt, xs, rel_tol=1e-13, abs_tol=1e-300,
                  max_iter=1000)
        return xs

    errs = []
    for f in faces:
        kp, km = k.copy(), k.copy()
        kp[f] += hstep
        km[f] -= hstep
        xp, xm = fd_solve(kp), fd_solve(km)
        fd = 0.5 * torch.dot(xp - xm, xp + xm - 2 * XT).item() / (2 * hstep)
        errs.append(abs(fd - g64[f]) / max(abs(g64[f]), 1e-3 * np.abs(g64).max()))
    H0.close()
    out["fd_max_rel_err_64_faces"] = float(max(errs))
    print(json.dumps(out), flush=True)


def hbm_ceiling(torch):
    """Best of 10: a 2 GiB FP64 sum (reads 2 GiB) and a 1 GiB copy (reads and
    writes 1 GiB each) with torch's own kernels, as TB/s of bytes moved."""
    dev = torch.device("cuda:0")
    n = 1 << 28
    x = torch.ones(n, dtype=torch.float64, device=dev)
    y = torch.empty(n // 2, dtype=torch.float64, device=dev)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    out = {}
    for name, fn, nbytes in (("read_sum", lambda: x.sum(), 8 * n),
                             ("copy", lambda: y.copy_(x[: n // 2]), 8 * n)):
        best = float("inf")
        for _ in range(11):
            e0.record()
            fn()
            e1.record()
            e1.synchronize()
            best = min(best, e0.elapsed_time(e1) / 1e3)
        out[name + "_tbs"] = round(nbytes / best / 1e12, 3)
    print(json.dumps({"metric": "HBM streaming bandwidth (best of 10)", **out,
                      "peak_tbs": HBM_PEAK_GBS / 1e3}), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the SA-AMG-PCG path here is built and "
              "measured on a single MI355X (no multi-GPU code); use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)
    import torch

    if args.hbm_ceiling:
        hbm_ceiling(torch)
        return
    if args.config != "c2":
        sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
        run_config(args)
        return

    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from linalglib_amd import lal, problems

    nx = args.nx
    rp, ci, v, m = problems.poisson3d(nx)
    nnz = int(rp[-1])
    xs = problems.uniform(m, seed=1)
    b = problems.rhs_from_solution(rp, ci, v, xs)

    torch.cuda.init()
    dev = torch.device("cuda:0")
    h = lal.Handle()
    RP = torch.from_numpy(rp).to(dev)
    CI = torch.from_numpy(ci).to(dev)
    V = torch.from_numpy(v).to(dev)
    B = torch.from_numpy(b).to(dev)
    torch.cuda.synchronize()

    if args.setup_only:
        # hierarchy builds alone, each timed between two synchronisations
        times = []
        for k in range(args.warmup + args.steps):
            ts = time.perf_counter()
            Hk = lal.Hierarchy(h, m, nnz, RP, CI, V)
            torch.cuda.synchronize()
            if k >= args.warmup:
                times.append(time.perf_counter() - ts)
            levels = Hk.levels
            Hk.close()
            torch.cuda.synchronize()
            print(f"# setup {k}: {(time.perf_counter() - ts) * 1e3:.2f} ms with destroy",
                  file=sys.stderr, flush=True)
        print(json.dumps({"metric": "saamg_setup_s", "workload": f"3D Poisson {nx}^3",
                          "n": m, "nnz": nnz, "levels": levels,
                          "setup_s": [round(t, 4) for t in times],
                          "setup_s_median": round(sorted(times)[len(times) // 2], 4)}),
              flush=True)
        return

    # ---------------------------------------------------------------- SpMV
    def spmv_loop(dtype, calls):
        vb = 8 if dtype == torch.float64 else 4
        fn = "lal_csrmv_f64" if vb == 8 else "lal_csrmv_f32"
        desc = lal.CsrDescr(h, m, m, nnz, RP, CI)
        copies = []
        for k in range(3):  # 3 rotating copies: > 256 MB Infinity Cache
            copies.append((RP.clone(), CI.clone(), V.to(dtype).clone(),
                           torch.rand(m, dtype=dtype, device=dev) * 2 - 1,
                           torch.zeros(m, dtype=dtype, device=dev)))
        L = lal.load()
        f = getattr(L, fn)
        argv = [(h.h, desc.d, m, m, nnz, 1.0, lal.ptr(c[0]), lal.ptr(c[1]), lal.ptr(c[2]),
                 lal.ptr(c[3]), 0.0, lal.ptr(c[4])) for c in copies]
        for k in range(4):
            f(*argv[k % 3])
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for k in range(calls):
            f(*argv[k % 3])
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 1e3 / calls
        # same loop on one copy (Infinity Cache resident)
        e0.record()
        for k in range(calls):
            f(*argv[0])
        e1.record()
        torch.cuda.synchronize()
        t_hot = e0.elapsed_time(e1) / 1e3 / calls
        byts = spmv_bytes(m, m, nnz, vb, vb)
        del copies
        return t, t_hot, byts

    t64, t64_hot, b64 = spmv_loop(torch.float64, args.spmv_calls)
    t32, t32_hot, b32 = spmv_loop(torch.float32, args.spmv_calls)
    print(f"# spmv fp64 {t64 * 1e6:.2f} us ({b64 / t64 / 1e9:.0f} GB/s, cache-resident "
          f"{b64 / t64_hot / 1e9:.0f}); fp32 {t32 * 1e6:.2f} us ({b32 / t32 / 1e9:.0f} GB/s)",
          file=sys.stderr, flush=True)
    if args.spmv_only:
        # calibration for hardware byte counters: a plain copy of 512 MiB
        src = torch.ones(64 << 20, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(f"# copy calibration: 5 x {src.numel() * 8} bytes read and written",
              file=sys.stderr, flush=True)
        return

    # ---------------------------------------------------------------- setup
    torch.cuda.synchronize()
    ts = time.perf_counter()
    H = lal.Hierarchy(h, m, nnz, RP, CI, V, build_f32=1)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - ts
    levels = []
    for lvl in range(H.levels + 1):
        lm, _, annz = H.dims(lvl, "A")
        pnnz = H.dims(lvl, "P")[2] if lvl < H.levels else 0
        cb = [H.column_bytes(lvl, k) if k == "A" or lvl < H.levels else 4 for k in "APR"]
        levels.append((lm, annz, pnnz, *cb))
    print(f"# setup {setup_s:.3f} s, levels {[(l[0], l[1]) for l in levels]}", file=sys.stderr,
          flush=True)

    # ---------------------------------------------------------------- solves
    X = torch.zeros(m, dtype=torch.float64, device=dev)

    def run(kind):
        X.zero_()
        if kind == "fp64":
            return H.pcg(B, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=1000, n1=1, n2=1)
        if kind == "vcycle":
            return H.pcg_mixed(B, X, rel_tol=1e-10, abs_tol=1e-300, max_iter=1000, n1=1, n2=1)
        if kind == "reliable":
            return H.pcg_reliable(B, X, delta=1e-2, rel_tol=1e-10, abs_tol=1e-300,
                                  max_iter=1000, n1=1, n2=1)
        return H.ir(B, X, rel_tol=1e-10, abs_tol=1e-300, n1=1, n2=1, inner_rel_tol=1e-4,
                    inner_max_iter=200, max_outer=30)

    results = {}
    for kind in ("fp64", "vcycle", "ir", "reliable"):
        for _ in range(args.warmup):
            run(kind)
        torch.cuda.synchronize()
        times = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            s0 = time.perf_counter()
            info = run(kind)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - s0)
        total = time.perf_counter() - t0
        x = X.cpu().numpy()
        r = b - (problems.rhs_from_solution(rp, ci, v, x))
        results[kind] = {
            "total_s": total,
            "median_s": statistics.median(times),
            "iters": info.iterations,
            "outer": info.outer_iterations,
            "converged": bool(info.converged),
            "rel_res_inf": float(np.abs(r).max() / np.abs(b).max()),
            "rel_err": float(np.linalg.norm(x - xs) / np.linalg.norm(xs)),
        }
        print(f"# {kind}: {results[kind]}", file=sys.stderr, flush=True)

    f64, mx = results["fp64"], results[args.mixed]
    solve64_gbs = solve_bytes(levels, f64["iters"], 8) / f64["median_s"] / 1e9
    rel = results["reliable"]
    rel_bytes = (solve_bytes(levels, rel["iters"], 4)
                 + rel["outer"] * reliable_update_bytes(levels[0][0], levels[0][1], levels[0][3]))
    solve_rel_gbs = rel_bytes / rel["median_s"] / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_threads()
        try:
            cpu = cpu_baseline(rp, ci, v, m, b, threads)
        except Exception as e:  # the baseline is context only
            cpu = {"value": None, "unit": "s", "cores": threads, "kind": "reference",
                   "sample": f"not run: {e}"}

    achieved = b64 / t64 / 1e9
    traffic = measured_traffic() if nx == 128 else None
    line = {
        "metric": "SA-AMG-PCG solve sec to 1e-10 & SpMV GB/s: mixed vs FP64, 1 MI355X",
        "value": round(mx["total_s"] / args.steps, 6),
        "unit": "s per solve (mixed FP32/FP64 SA-AMG-PCG to rel. residual 1e-10)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mx["total_s"] / args.steps * 1e3, 4),
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {
            "workload": f"3D 7-point Poisson {nx}^3 SA-AMG-PCG to 1e-10 (1+1 Jacobi V-cycle)",
            "n": m, "nnz": nnz, "levels": H.levels,
            "level_sizes": [[l[0], l[1]] for l in levels],
            "mixed_scheme": {"vcycle": "FP64 PCG, FP32 V-cycle",
                             "ir": "FP32 PCG in FP64 refinement",
                             "reliable": "FP32 PCG and V-cycle, FP64 reliable updates"}[
                                 args.mixed],
            "parallelism": "single GPU",
        },
        "fp64_solve_s": round(f64["total_s"] / args.steps, 6),
        "mixed_solve_s": round(mx["total_s"] / args.steps, 6),
        "speedup_mixed_vs_fp64": round(f64["total_s"] / mx["total_s"], 3),
        "fp64_iterations": f64["iters"],
        "mixed_iterations": mx["iters"],
        "mixed_outer_iterations": mx["outer"],
        "mixed_vcycle_solve_s": round(results["vcycle"]["total_s"] / args.steps, 6),
        "mixed_vcycle_iterations": results["vcycle"]["iters"],
        "mixed_ir_solve_s": round(results["ir"]["total_s"] / args.steps, 6),
        "mixed_ir_inner_iterations": results["ir"]["iters"],
        "mixed_ir_outer_iterations": results["ir"]["outer"],
        "mixed_reliable_solve_s": round(results["reliable"]["total_s"] / args.steps, 6),
        "mixed_reliable_iterations": results["reliable"]["iters"],
        "mixed_reliable_updates": results["reliable"]["outer"],
        "mixed_reliable_rel_residual_inf": results["reliable"]["rel_res_inf"],
        "fp64_rel_residual_inf": f64["rel_res_inf"],
        "mixed_rel_residual_inf": mx["rel_res_inf"],
        "fp64_rel_err_vs_xstar": f64["rel_err"],
        "mixed_rel_err_vs_xstar": mx["rel_err"],
        "setup_s": round(setup_s, 4),
        "spmv_fp64_gbs": round(achieved, 1),
        "spmv_fp32_gbs": round(b32 / t32 / 1e9, 1),
        "spmv_fp64_cache_resident_gbs": round(b64 / t64_hot / 1e9, 1),
        "spmv_fp64_us": round(t64 * 1e6, 3),
        "solve_fp64_algorithmic_gbs": round(solve64_gbs, 1),
        "solve_fp64_roofline_frac": round(solve64_gbs / HBM_PEAK_GBS, 4),
        "solve_mixed_reliable_algorithmic_gbs": round(solve_rel_gbs, 1),
        "solve_mixed_reliable_roofline_frac": round(solve_rel_gbs / HBM_PEAK_GBS, 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "quad_kernel<double, 0, false> (FP64 CSR SpMV, 3 rotating copies)",
            "bytes_per_launch": b64,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
