"""ctypes binding of the C ABI in include/lal.h, with torch tensors as the
device buffers.

torch is plumbing here: it allocates HBM and provides the stream.  Every
handle is bound to torch's current stream, so library kernels are ordered
with torch's own copies.  If liblal.so is missing, or the device is not
gfx950, calls raise LalError -- there is no CPU fallback.
"""
import ctypes as C
import os

from . import build as _build

_lib = None

LAL_OK = 0
STATUS = {
    0: "ok", 1: "invalid argument", 2: "out of device memory", 3: "HIP runtime error",
    4: "not converged", 5: "NaN in residual", 6: "no usable device", 7: "unsupported input",
}


class LalError(RuntimeError):
    def __init__(self, status, where, detail=""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)}"
                         + (f" ({detail})" if detail else ""))


vp = C.c_void_p
i32 = C.c_int
f64 = C.c_double
f32 = C.c_float
sz = C.c_size_t


class SaamgOptions(C.Structure):
    _fields_ = [("max_levels", i32), ("eps", f64), ("relax", f64), ("coarse_solver", i32),
                ("exact", i32), ("build_f32", i32)]


class SolveOptions(C.Structure):
    _fields_ = [("rel_tol", f64), ("abs_tol", f64), ("max_iter", i32), ("n1", i32),
                ("n2", i32), ("inner_rel_tol", f64), ("inner_max_iter", i32),
                ("max_outer", i32), ("smoother", i32)]


class SolveInfo(C.Structure):
    _fields_ = [("iterations", i32), ("outer_iterations", i32), ("converged", i32),
                ("initial_res_norm", f64), ("final_res_norm", f64)]


# name -> argtypes (all return int status unless listed in _RESTYPES)
_SIGS = {
    "lal_create": [C.POINTER(vp)],
    "lal_destroy": [vp],
    "lal_set_stream": [vp, vp],
    "lal_get_stream": [vp, C.POINTER(vp)],
    "lal_sync": [vp],
    "lal_alloc": [C.POINTER(vp), sz],
    "lal_free": [vp],
    "lal_memcpy_h2d": [vp, vp, vp, sz],
    "lal_memcpy_d2h": [vp, vp, vp, sz],
    "lal_memcpy_d2d": [vp, vp, vp, sz],
    "lal_fill_f64": [vp, vp, sz, f64],
    "lal_fill_f32": [vp, vp, sz, f32],
    "lal_fill_i32": [vp, vp, sz, C.c_int32],
    "lal_fill_i64": [vp, vp, sz, C.c_int64],
    "lal_convert_f64_f32": [vp, vp, vp, sz],
    "lal_convert_f32_f64": [vp, vp, vp, sz],
    "lal_axpy_f64": [vp, i32, f64, vp, vp],
    "lal_axpby_f64": [vp, i32, f64, vp, f64, vp],
    "lal_axpbypgz_f64": [vp, i32, f64, vp, f64, vp, f64, vp],
    "lal_axpy_f32": [vp, i32, f32, vp, vp],
    "lal_axpby_f32": [vp, i32, f32, vp, f32, vp],
    "lal_dot_f64": [vp, i32, vp, vp, C.POINTER(f64)],
    "lal_axpy_dot_f64": [vp, i32, f64, vp, vp, vp, C.POINTER(f64)],
    "lal_residual_f32": [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp],
    "lal_transpose_f32": [vp, i32, i32, i32, vp, vp, vp, vp, vp, vp],
    "lal_pmis_init": [vp, i32, i32, vp, vp, vp, vp],
    "lal_pmis_round": [vp, i32, i32, vp, vp, vp, vp, vp, vp, vp, C.POINTER(i32)],
    "lal_pmis_join": [vp, i32, i32, vp, vp, vp, vp, vp, vp],
    "lal_dot_f32": [vp, i32, vp, vp, C.POINTER(f64)],
    "lal_norm_inf_f64": [vp, i32, vp, C.POINTER(f64)],
    "lal_norm_inf_f32": [vp, i32, vp, C.POINTER(f64)],
    "lal_exclusive_scan_i32": [vp, i32, vp],
    "lal_exclusive_scan_i64": [vp, i32, vp],
    "lal_exclusive_scan_i64_noalloc": [vp, i32, vp],
    "lal_csrmv_analysis": [vp, i32, i32, i32, vp, vp, i32, C.POINTER(vp)],
    "lal_csrmv_descr_destroy": [vp],
    "lal_csrmv_f64": [vp, vp, i32, i32, i32, f64, vp, vp, vp, vp, f64, vp],
    "lal_csrmv_f32": [vp, vp, i32, i32, i32, f32, vp, vp, vp, vp, f32, vp],
    "lal_residual_f64": [vp, vp, i32, i32, i32, vp, vp, vp, vp, vp, vp],
    "lal_diagonal_f64": [vp, i32, vp, vp, vp, vp],
    "lal_transpose_f64": [vp, i32, i32, i32, vp, vp, vp, vp, vp, vp],
    "lal_csrgemm_nnz": [vp, i32, i32, i32, i32, vp, vp, i32, vp, vp, vp, C.POINTER(i32)],
    "lal_csrgemm_solve_f64": [vp, i32, i32, i32, i32, vp, vp, vp, i32, vp, vp, vp, vp, vp, vp],
    "lal_sa_strength_f64": [vp, i32, i32, vp, vp, vp, f64, vp],
    "lal_pmis_aggregate": [vp, i32, i32, vp, vp, vp, vp],
    "lal_sa_prolongator_nnz": [vp, i32, i32, vp, vp, vp, vp, vp, C.POINTER(i32),
                               C.POINTER(i32)],
    "lal_sa_prolongator_fill_f64": [vp, i32, i32, vp, vp, vp, vp, vp, f64, vp, vp, vp],
    "lal_saamg_default_options": [C.POINTER(SaamgOptions)],
    "lal_saamg_setup": [vp, i32, i32, i32, vp, vp, vp, C.POINTER(SaamgOptions), C.POINTER(vp)],
    "lal_uaamg_setup": [vp, i32, i32, i32, vp, vp, vp, C.POINTER(SaamgOptions), C.POINTER(vp)],
    "lal_uaamg_default_options": [C.POINTER(SaamgOptions)],
    "lal_hierarchy_destroy": [vp],
    "lal_hierarchy_levels": [vp, C.POINTER(i32)],
    "lal_hierarchy_level": [vp, i32, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                            C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)],
    "lal_hierarchy_column_bytes": [vp, i32, i32, C.POINTER(i32)],
    "lal_amg_precond_f64": [vp, vp, vp, vp, i32, i32],
    "lal_solve_default_options": [C.POINTER(SolveOptions)],
    "lal_pcg_solve_f64": [vp, vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_pcg_solve_op_f64": [vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, C.POINTER(SolveOptions),
                             C.POINTER(SolveInfo)],
    "lal_ir_solve": [vp, vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_pcg_solve_mixed": [vp, vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_pcg_solve_reliable": [vp, vp, vp, vp, C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_amg_solve": [vp, vp, vp, vp, i32, i32, i32, i32, C.POINTER(SolveOptions),
                      C.POINTER(SolveInfo)],
    "lal_sddmm_f64": [vp, i32, i32, vp, vp, f64, vp, vp, vp],
    "lal_jacobi_solve_f64": [vp, i32, vp, vp, vp],
    "lal_bicgstab_solve": [vp, vp, i32, i32, vp, vp, vp, vp, vp, i32, vp,
                           C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_gmres_solve": [vp, vp, i32, i32, vp, vp, vp, vp, vp, i32, i32, vp,
                        C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_adjoint_solve": [vp, vp, vp, vp, i32, C.POINTER(SolveOptions), C.POINTER(SolveInfo)],
    "lal_status_string": [i32],
    "lal_last_error": [],
}
_RESTYPES = {"lal_status_string": C.c_char_p, "lal_last_error": C.c_char_p,
             "lal_saamg_default_options": None,
             "lal_uaamg_default_options": None, "lal_solve_default_options": None}


def lib_path():
    return _build.LIB


def load():
    """Loads liblal.so; raises LalError when it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    path = lib_path()
    if not os.path.exists(path):
        raise LalError(6, "load", f"{path} not built: run python -m linalglib_amd.build")
    L = C.CDLL(path)
    for name, args in _SIGS.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = L
    return L


def check(status, where):
    if status != LAL_OK:
        detail = ""
        try:
            detail = load().lal_last_error().decode()
        except Exception:
            pass
        raise LalError(status, where, detail)


def call(name, *args):
    L = load()
    check(getattr(L, name)(*args), name)


def ptr(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    return C.c_void_p(t.data_ptr())


class Handle:
    """A lal context bound to torch's current stream on the current device."""

    def __init__(self):
        import torch

        if not torch.cuda.is_available():
            raise LalError(6, "Handle", "no GPU visible to torch")
        L = load()
        h = vp()
        check(L.lal_create(C.byref(h)), "lal_create")
        self.h = h
        self.L = L
        check(L.lal_set_stream(h, C.c_void_p(torch.cuda.current_stream().cuda_stream)),
              "lal_set_stream")

    def close(self):
        if getattr(self, "h", None):
            self.L.lal_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(self.L.lal_sync(self.h), "lal_sync")


def d2h(handle, dptr, count, dtype):
    """Copies `count` elements at a raw device pointer into a numpy array."""
    import numpy as np

    out = np.empty(max(int(count), 0), dtype=dtype)
    if count:
        check(load().lal_memcpy_d2h(handle.h, C.c_void_p(out.ctypes.data), dptr,
                                    out.nbytes), "lal_memcpy_d2h")
    return out


def saamg_options(**kw):
    o = SaamgOptions()
    load().lal_saamg_default_options(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


SMOOTHER_CODES = {"jacobi": 0, "gs": 1, "sgs": 2, "sor": 3, "ssor": 4, "mcgs": 5, "mcsgs": 6}


def solve_options(**kw):
    o = SolveOptions()
    load().lal_solve_default_options(C.byref(o))
    for k, v in kw.items():
        if k == "smoother" and isinstance(v, str):
            v = SMOOTHER_CODES[v]
        setattr(o, k, v)
    return o


class Hierarchy:
    """SA-AMG hierarchy built on the device from device CSR arrays (torch
    tensors int32/int32/float64)."""

    KIND = {"A": 0, "P": 1, "R": 2}

    def __init__(self, handle, m, nnz, row_ptr, col_ind, val, unsmoothed=False, **opts):
        """unsmoothed=True builds LinAlgLib's uaamg_setup hierarchy (eps from
        0.001, one 1.0 per prolongator row) instead of saamg_setup's."""
        self.handle = handle
        self.opts = saamg_options(**opts)
        if unsmoothed and "eps" not in opts:
            self.opts.eps = 0.001
        fn = "lal_uaamg_setup" if unsmoothed else "lal_saamg_setup"
        hp = vp()
        check(getattr(load(), fn)(handle.h, m, m, nnz, ptr(row_ptr), ptr(col_ind), ptr(val),
                                  C.byref(self.opts), C.byref(hp)), fn)
        self.hp = hp
        self.m = m

    def close(self):
        if getattr(self, "hp", None):
            load().lal_hierarchy_destroy(self.hp)
            self.hp = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def levels(self):
        n = i32()
        check(load().lal_hierarchy_levels(self.hp, C.byref(n)), "lal_hierarchy_levels")
        return n.value

    def dims(self, level, kind):
        """(m, n, nnz) of A/P/R at a level, without copying the arrays."""
        m, n, nnz = i32(), i32(), i32()
        rp, ci, v = vp(), vp(), vp()
        check(load().lal_hierarchy_level(self.hp, level, self.KIND[kind], C.byref(m), C.byref(n),
                                         C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(v)),
              "lal_hierarchy_level")
        return m.value, n.value, nnz.value

    def column_bytes(self, level, kind):
        """Bytes per column index the fast path reads for A/P/R at a level
        (2: 16-bit offsets from each row block's base, 4: int32)."""
        out = i32()
        check(load().lal_hierarchy_column_bytes(self.hp, level, self.KIND[kind], C.byref(out)),
              "lal_hierarchy_column_bytes")
        return out.value

    def level(self, level, kind):
        """Host copy (row_ptr, col_ind, val, m, n) of A/P/R at a level."""
        import numpy as np

        m, n, nnz = i32(), i32(), i32()
        rp, ci, v = vp(), vp(), vp()
        check(load().lal_hierarchy_level(self.hp, level, self.KIND[kind], C.byref(m), C.byref(n),
                                         C.byref(nnz), C.byref(rp), C.byref(ci), C.byref(v)),
              "lal_hierarchy_level")
        return (d2h(self.handle, rp, m.value + 1, np.int32),
                d2h(self.handle, ci, nnz.value, np.int32),
                d2h(self.handle, v, nnz.value, np.float64), m.value, n.value)

    def precond(self, r, z, n1=1, n2=1):
        call("lal_amg_precond_f64", self.handle.h, self.hp, ptr(r), ptr(z), n1, n2)

    def pcg(self, b, x, **opts):
        o = solve_options(**opts)
        info = SolveInfo()
        call("lal_pcg_solve_f64", self.handle.h, self.hp, ptr(b), ptr(x), C.byref(o),
             C.byref(info))
        return info

    def pcg_op(self, A, b, x, **opts):
        """PCG whose Krylov products use A = (row_ptr, col_ind, val, m, nnz,
        descr or None), a matrix of this hierarchy's size, preconditioned by
        one V-cycle of this hierarchy (cg_solver::solve(A, x, b, precond))."""
        rp, ci, v, m, nnz, d = A
        o = solve_options(**opts)
        info = SolveInfo()
        call("lal_pcg_solve_op_f64", self.handle.h, d.d if d is not None else None, m, nnz,
             ptr(rp), ptr(ci), ptr(v), self.hp, ptr(b), ptr(x), C.byref(o), C.byref(info))
        return info

    def ir(self, b, x, **opts):
        o = solve_options(**opts)
        info = SolveInfo()
        call("lal_ir_solve", self.handle.h, self.hp, ptr(b), ptr(x), C.byref(o), C.byref(info))
        return info

    CYCLES = {"v": 0, "w": 1, "f": 2}
    SMOOTHERS = SMOOTHER_CODES

    def amg_solve(self, b, x, n1=1, n2=1, cycle="v", smoother="jacobi", **opts):
        """AMG as a solver (x in place); max_iter caps the number of cycles."""
        o = solve_options(**opts)
        info = SolveInfo()
        call("lal_amg_solve", self.handle.h, self.hp, ptr(b), ptr(x), n1, n2,
             self.CYCLES[cycle], self.SMOOTHERS[smoother], C.byref(o), C.byref(info))
        return info

    def pcg_reliable(self, b, x, delta=1e-2, **opts):
        """FP32 PCG (FP32 operator, vectors and V-cycle) with FP64 reliable
        updates every `delta` of residual reduction (needs build_f32);
        info.outer_iterations counts the updates."""
        o = solve_options(**opts)
        o.inner_rel_tol = delta
        info = SolveInfo()
        call("lal_pcg_solve_reliable", self.handle.h, self.hp, ptr(b), ptr(x), C.byref(o),
             C.byref(info))
        return info

    def pcg_mixed(self, b, x, **opts):
        """FP64 PCG with the FP32 V-cycle as preconditioner (needs build_f32)."""
        o = solve_options(**opts)
        info = SolveInfo()
        call("lal_pcg_solve_mixed", self.handle.h, self.hp, ptr(b), ptr(x), C.byref(o),
             C.byref(info))
        return info


class CsrDescr:
    """SpMV analysis descriptor (row blocks) for one matrix pattern."""

    def __init__(self, h, m, n, nnz, row_ptr, col_ind, exact=False):
        d = vp()
        check(load().lal_csrmv_analysis(h.h, m, n, nnz, ptr(row_ptr), ptr(col_ind),
                                        int(exact), C.byref(d)), "lal_csrmv_analysis")
        self.d = d

    def __del__(self):
        if getattr(self, "d", None):
            load().lal_csrmv_descr_destroy(self.d)
            self.d = None


PRECONDS = {"none": 0, "jacobi": 1, "saamg": 2}


def _krylov_args(A, precond, hier):
    if precond not in PRECONDS:
        raise ValueError(f"precond must be one of {sorted(PRECONDS)}")
    if precond == "saamg" and hier is None:
        raise ValueError("precond='saamg' needs the Hierarchy built for A")
    rp, ci, val, m, nnz, descr = A
    return (ptr(rp), ptr(ci), ptr(val), m, nnz, descr.d if descr is not None else None,
            PRECONDS[precond], hier.hp if (hier is not None and precond == "saamg") else None)


def jacobi_solve(handle, rhs, diag, x):
    """x = rhs / diag elementwise (LinAlgLib's jacobi_precond::solve)."""
    call("lal_jacobi_solve_f64", handle.h, rhs.numel(), ptr(rhs), ptr(diag), ptr(x))


def bicgstab(handle, A, b, x, precond="none", hier=None, **opts):
    """BiCGSTAB on A = (row_ptr, col_ind, val, m, nnz, CsrDescr or None); x is
    the initial guess and the result.  precond: "none", "jacobi" or "saamg"
    (one V-cycle of `hier`, opts n1/n2 sweeps)."""
    rp, ci, val, m, nnz, d, kind, hp = _krylov_args(A, precond, hier)
    o = solve_options(**opts)
    info = SolveInfo()
    call("lal_bicgstab_solve", handle.h, d, m, nnz, rp, ci, val, ptr(b), ptr(x), kind, hp,
         C.byref(o), C.byref(info))
    return info


def gmres(handle, A, b, x, restart=30, precond="none", hier=None, **opts):
    """Restarted, left-preconditioned GMRES(restart); arguments as bicgstab."""
    rp, ci, val, m, nnz, d, kind, hp = _krylov_args(A, precond, hier)
    o = solve_options(**opts)
    info = SolveInfo()
    call("lal_gmres_solve", handle.h, d, m, nnz, rp, ci, val, ptr(b), ptr(x), int(restart), kind,
         hp, C.byref(o), C.byref(info))
    return info
