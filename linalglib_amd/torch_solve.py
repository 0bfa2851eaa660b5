"""Differentiable SA-AMG-PCG solve for PyTorch (float64 tensors on cuda:0).

    x = saamg_solve(val, b, row_ptr, col_ind)

solves A x = b for the SPD matrix A with CSR pattern (row_ptr, col_ind) and
values `val`.  Backward, for a loss L with g = dL/dx:

    dL/db    = lambda, the solution of A lambda = g (the adjoint system; A is
               symmetric, so the forward hierarchy preconditions it too)
    dL/dA_ij = -lambda_i x_j over A's pattern (lal_sddmm_f64)

Gradients with respect to parameters that assemble `val` (face
conductances, edge weights) follow from torch's own autograd through the
assembly, e.g. `val = torch.zeros(nnz).index_add(0, entry, sign * k[face])`.
mixed selects the precision of both solves: False (FP64 throughout), "vcycle"
(FP64 PCG with the V-cycle on FP32 levels; also what True means), "ir"
(FP32 inner SA-AMG-PCG inside FP64 iterative refinement) or "reliable" (FP32
PCG and V-cycle with FP64 reliable updates every 1e-2 of residual reduction).
"""
import ctypes as C

import torch

from . import lal


class _Solve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, val, b, row_ptr, col_ind, handle, mixed, tol, hier, max_iter):
        m = b.numel()
        nnz = val.numel()
        val_c = val.detach().contiguous()
        x = torch.zeros(m, dtype=torch.float64, device=b.device)
        b_c = b.detach().contiguous()
        opts = dict(rel_tol=tol, abs_tol=1e-300, max_iter=max_iter)
        ctx.op = None
        if hier is not None:
            # reuse a hierarchy built earlier (for this A or one of the same
            # size): PCG on the current values, its V-cycle as preconditioner
            H = hier
            ctx.op = (row_ptr, col_ind, val_c, m, nnz, None)
            info = H.pcg_op(ctx.op, b_c, x, **opts)
        else:
            H = lal.Hierarchy(handle, m, nnz, row_ptr, col_ind, val_c,
                              build_f32=int(mixed != 0))
            solve = {0: H.pcg, 1: H.ir, 2: H.pcg_mixed, 3: H.pcg_reliable}[mixed]
            info = solve(b_c, x, **opts)
        if not info.converged:
            raise lal.LalError(4, "saamg_solve forward",
                               f"residual {info.final_res_norm:.3e} after {info.iterations}")
        ctx.H = H
        ctx.handle = handle
        ctx.mixed = mixed
        ctx.tol = tol
        ctx.max_iter = max_iter
        ctx.keep = (row_ptr, col_ind, val_c)
        ctx.save_for_backward(x)
        ctx.info = info
        return x

    @staticmethod
    def backward(ctx, gx):
        (x,) = ctx.saved_tensors
        row_ptr, col_ind, val_c = ctx.keep
        g = gx.detach().contiguous().to(torch.float64)
        lam = torch.empty_like(g)
        o = lal.solve_options(rel_tol=ctx.tol, abs_tol=1e-300, max_iter=ctx.max_iter)
        if ctx.mixed == 3:
            o.inner_rel_tol = 1e-2  # the reliable-update threshold
        info = lal.SolveInfo()
        if ctx.op is not None:
            # the Krylov products on the current A, the V-cycle of the
            # caller's hierarchy (built for some A of the same pattern)
            lam.zero_()
            info = ctx.H.pcg_op(ctx.op, g, lam, rel_tol=ctx.tol, abs_tol=1e-300,
                                max_iter=ctx.max_iter)
        else:
            lal.call("lal_adjoint_solve", ctx.handle.h, ctx.H.hp, lal.ptr(g), lal.ptr(lam),
                     int(ctx.mixed), C.byref(o), C.byref(info))
        if not info.converged:
            raise lal.LalError(4, "saamg_solve backward (adjoint solve)",
                               f"residual {info.final_res_norm:.3e} after {info.iterations}")
        dval = None
        if ctx.needs_input_grad[0]:
            dval = torch.empty(val_c.numel(), dtype=torch.float64, device=x.device)
            lal.call("lal_sddmm_f64", ctx.handle.h, x.numel(), val_c.numel(),
                     lal.ptr(row_ptr), lal.ptr(col_ind), -1.0, lal.ptr(lam), lal.ptr(x),
                     lal.ptr(dval))
        db = lam if ctx.needs_input_grad[1] else None
        return dval, db, None, None, None, None, None, None, None


_handles = {}


def _handle():
    dev = torch.cuda.current_device()
    stream = torch.cuda.current_stream().cuda_stream
    key = (dev, stream)
    if key not in _handles:
        _handles[key] = lal.Handle()
    return _handles[key]


def saamg_solve(val, b, row_ptr, col_ind, mixed=False, tol=1e-12, hierarchy=None,
                max_iter=1000):
    """x = A^-1 b with A = CSR(row_ptr, col_ind, val); differentiable in val and b.

    hierarchy: a lal.Hierarchy built earlier for a matrix of the same size
    (e.g. the first of a sequence of optimisation steps).  The forward and
    adjoint solves then run FP64 PCG on the current values with that
    hierarchy's V-cycle as the preconditioner instead of building a new one
    (mixed must be False).  A solve that does not converge within max_iter
    iterations raises LalError, in the forward and in the adjoint pass."""
    if val.dtype != torch.float64 or b.dtype != torch.float64:
        raise TypeError("saamg_solve takes float64 values and right-hand side")
    modes = {False: 0, True: 2, "ir": 1, "vcycle": 2, "reliable": 3}
    if mixed not in modes:
        raise ValueError(f"mixed must be False, True, 'ir', 'vcycle' or 'reliable', "
                         f"not {mixed!r}")
    if hierarchy is not None and mixed is not False:
        raise ValueError("a reused hierarchy runs the FP64 PCG only (mixed=False)")
    handle = hierarchy.handle if hierarchy is not None else _handle()
    return _Solve.apply(val, b, row_ptr, col_ind, handle, modes[mixed], float(tol), hierarchy,
                        int(max_iter))
