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
    def forward(ctx, val, b, row_ptr, col_ind, handle, mixed, tol):
        m = b.numel()
        nnz = val.numel()
        val_c = val.detach().contiguous()
        H = lal.Hierarchy(handle, m, nnz, row_ptr, col_ind, val_c, build_f32=int(mixed != 0))
        x = torch.zeros(m, dtype=torch.float64, device=b.device)
        b_c = b.detach().contiguous()
        opts = dict(rel_tol=tol, abs_tol=1e-300, max_iter=1000)
        solve = {0: H.pcg, 1: H.ir, 2: H.pcg_mixed, 3: H.pcg_reliable}[mixed]
        info = solve(b_c, x, **opts)
        if not info.converged:
            raise lal.LalError(4, "saamg_solve forward",
                               f"residual {info.final_res_norm:.3e} after {info.iterations}")
        ctx.H = H
        ctx.handle = handle
        ctx.mixed = mixed
        ctx.tol = tol
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
        o = lal.solve_options(rel_tol=ctx.tol, abs_tol=1e-300, max_iter=1000)
        if ctx.mixed == 3:
            o.inner_rel_tol = 1e-2  # the reliable-update threshold
        info = lal.SolveInfo()
        lal.call("lal_adjoint_solve", ctx.handle.h, ctx.H.hp, lal.ptr(g), lal.ptr(lam),
                 int(ctx.mixed), C.byref(o), C.byref(info))
        dval = None
        if ctx.needs_input_grad[0]:
            dval = torch.empty(val_c.numel(), dtype=torch.float64, device=x.device)
            lal.call("lal_sddmm_f64", ctx.handle.h, x.numel(), val_c.numel(),
                     lal.ptr(row_ptr), lal.ptr(col_ind), -1.0, lal.ptr(lam), lal.ptr(x),
                     lal.ptr(dval))
        db = lam if ctx.needs_input_grad[1] else None
        return dval, db, None, None, None, None, None


_handles = {}


def _handle():
    dev = torch.cuda.current_device()
    stream = torch.cuda.current_stream().cuda_stream
    key = (dev, stream)
    if key not in _handles:
        _handles[key] = lal.Handle()
    return _handles[key]


def saamg_solve(val, b, row_ptr, col_ind, mixed=False, tol=1e-12):
    """x = A^-1 b with A = CSR(row_ptr, col_ind, val); differentiable in val and b."""
    if val.dtype != torch.float64 or b.dtype != torch.float64:
        raise TypeError("saamg_solve takes float64 values and right-hand side")
    modes = {False: 0, True: 2, "ir": 1, "vcycle": 2, "reliable": 3}
    if mixed not in modes:
        raise ValueError(f"mixed must be False, True, 'ir', 'vcycle' or 'reliable', "
                         f"not {mixed!r}")
    return _Solve.apply(val, b, row_ptr, col_ind, _handle(), modes[mixed], float(tol))
