"""ctypes access to the CPU oracle (liboracle.so) and, where it was built, to
LinAlgLib's own host library (oracle/_ref/libref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.  The product path never
imports this module.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libref.so")

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_ip = C.POINTER(C.c_int)
_dp = C.POINTER(C.c_double)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class CSR:
    """Host CSR triple with int32 indices and float64 values."""

    def __init__(self, row_ptr, col_ind, val, m, n):
        self.row_ptr = _i32(row_ptr)
        self.col_ind = _i32(col_ind)
        self.val = _f64(val)
        self.m = int(m)
        self.n = int(n)

    @property
    def nnz(self):
        return int(self.row_ptr[-1]) if self.m >= 0 else 0

    def to_scipy(self):
        import scipy.sparse as sp

        return sp.csr_matrix((self.val, self.col_ind, self.row_ptr), shape=(self.m, self.n))


class _CsrStruct(C.Structure):
    _fields_ = [
        ("m", C.c_int),
        ("n", C.c_int),
        ("nnz", C.c_int),
        ("row_ptr", _ip),
        ("col_ind", _ip),
        ("val", _dp),
    ]


def _struct_to_csr(s):
    m, n, nnz = s.m, s.n, s.nnz
    rp = np.ctypeslib.as_array(s.row_ptr, shape=(m + 1,)).copy()
    ci = np.ctypeslib.as_array(s.col_ind, shape=(max(nnz, 1),))[:nnz].copy()
    v = np.ctypeslib.as_array(s.val, shape=(max(nnz, 1),))[:nnz].copy()
    return CSR(rp, ci, v, m, n)


def _struct_from(csr):
    s = _CsrStruct()
    s.m, s.n, s.nnz = csr.m, csr.n, csr.nnz
    s.row_ptr = csr.row_ptr.ctypes.data_as(_ip)
    s.col_ind = csr.col_ind.ctypes.data_as(_ip)
    s.val = csr.val.ctypes.data_as(_dp)
    return s


class Oracle:
    """The C restatement of the reference host path."""

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle)")
        L = C.CDLL(path)
        self.L = L
        L.or_set_threads.argtypes = [C.c_int]
        L.or_csrmv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _i32p, _i32p, _f64p,
                               _f64p, C.c_double, _f64p]
        L.or_residual.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.or_dot.argtypes = [C.c_int, _f64p, _f64p]
        L.or_dot.restype = C.c_double
        L.or_norm_inf.argtypes = [C.c_int, _f64p]
        L.or_norm_inf.restype = C.c_double
        L.or_axpy.argtypes = [C.c_int, C.c_double, _f64p, _f64p]
        L.or_axpby.argtypes = [C.c_int, C.c_double, _f64p, C.c_double, _f64p]
        L.or_diagonal.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p]
        L.or_exclusive_scan_i64.argtypes = [C.c_int, _i64p]
        L.or_strength.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_double, _i32p]
        L.or_hash1.argtypes = [C.c_uint]
        L.or_hash1.restype = C.c_uint
        L.or_pmis_init.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p]
        L.or_pmis_round.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p, _i64p, _i32p,
                                    C.POINTER(C.c_int)]
        L.or_pmis_join.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i32p, _i64p, _i32p]
        L.or_aggregates_pmis.argtypes = [C.c_int, _i32p, _i32p, _i32p, _i64p, C.c_void_p]
        L.or_prolongator.argtypes = [C.c_int, _i32p, _i32p, _f64p, _i32p, _i64p, C.c_double,
                                     C.POINTER(_CsrStruct)]
        L.or_transpose.argtypes = [C.POINTER(_CsrStruct), C.POINTER(_CsrStruct)]
        L.or_csrgemm.argtypes = [C.POINTER(_CsrStruct)] * 3
        L.or_csr_free.argtypes = [C.POINTER(_CsrStruct)]
        L.or_saamg_setup.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int,
                                     C.c_double, C.c_double]
        L.or_saamg_setup.restype = C.c_void_p
        L.or_uaamg_setup.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int,
                                     C.c_double]
        L.or_uaamg_setup.restype = C.c_void_p
        L.or_hierarchy_free.argtypes = [C.c_void_p]
        L.or_hierarchy_levels.argtypes = [C.c_void_p]
        L.or_hierarchy_dims.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip]
        L.or_hierarchy_copy.argtypes = [C.c_void_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.or_cholesky_factor.argtypes = [C.c_int, _f64p]
        L.or_jacobi_iteration.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, _f64p]
        L.or_amg_precond.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.c_int, C.c_int]
        L.or_pcg_saamg_op.argtypes = [C.c_int, _i32p, _i32p, _f64p, C.c_void_p, _f64p, _f64p,
                                      C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                      C.c_int, _dp]
        L.or_pcg_saamg.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                   C.c_int, C.c_int, C.c_int, _dp]
        L.or_cg.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                            C.c_double, C.c_int, _dp]
        L.or_smooth.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.or_bicgstab.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int,
                                  C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.or_bicgstab.restype = C.c_int
        L.or_gmres.argtypes = [C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int, C.c_int,
                               C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.or_gmres.restype = C.c_int
        L.or_amg_solve.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_double, C.c_double, C.c_int, _dp]
        L.or_amg_solve.restype = C.c_int

    def set_threads(self, n):
        self.L.or_set_threads(int(n))

    # ---- BLAS-1 / SpMV ----
    def csrmv(self, A, x, alpha=1.0, beta=0.0, y=None):
        y = np.zeros(A.m) if y is None else _f64(y).copy()
        self.L.or_csrmv(A.m, A.n, A.nnz, alpha, A.row_ptr, A.col_ind, A.val, _f64(x), beta, y)
        return y

    def residual(self, A, x, b):
        r = np.zeros(A.m)
        self.L.or_residual(A.m, A.row_ptr, A.col_ind, A.val, _f64(x), _f64(b), r)
        return r

    def dot(self, x, y):
        return self.L.or_dot(len(x), _f64(x), _f64(y))

    def norm_inf(self, x):
        return self.L.or_norm_inf(len(x), _f64(x))

    def axpy(self, alpha, x, y):
        y = _f64(y).copy()
        self.L.or_axpy(len(x), alpha, _f64(x), y)
        return y

    def axpby(self, alpha, x, beta, y):
        y = _f64(y).copy()
        self.L.or_axpby(len(x), alpha, _f64(x), beta, y)
        return y

    def diagonal(self, A):
        d = np.zeros(A.m)
        self.L.or_diagonal(A.m, A.row_ptr, A.col_ind, A.val, d)
        return d

    def exclusive_scan_i64(self, x):
        x = np.ascontiguousarray(x, dtype=np.int64).copy()
        self.L.or_exclusive_scan_i64(len(x), x)
        return x

    # ---- setup pieces ----
    def strength(self, A, eps):
        conn = np.zeros(A.nnz, dtype=np.int32)
        self.L.or_strength(A.m, A.row_ptr, A.col_ind, A.val, eps, conn)
        return conn

    def hash1(self, x):
        return self.L.or_hash1(int(x))

    def pmis_init(self, A, conn):
        state = np.zeros(A.m, dtype=np.int32)
        h = np.zeros(A.m, dtype=np.int32)
        self.L.or_pmis_init(A.m, A.row_ptr, _i32(conn), state, h)
        return state, h

    def pmis_round(self, A, conn, state, h, aggregates, max_state):
        agg = np.ascontiguousarray(aggregates, dtype=np.int64).copy()
        ms = _i32(max_state).copy()
        complete = C.c_int(1)
        self.L.or_pmis_round(A.m, A.row_ptr, A.col_ind, _i32(conn), _i32(state), _i32(h), agg,
                             ms, C.byref(complete))
        return agg, ms, bool(complete.value)

    def aggregates(self, A, conn):
        agg = np.zeros(A.m, dtype=np.int64)
        self.L.or_aggregates_pmis(A.m, A.row_ptr, A.col_ind, _i32(conn), agg, None)
        return agg

    def prolongator(self, A, conn, aggregates, relax=2.0 / 3.0):
        s = _CsrStruct()
        self.L.or_prolongator(A.m, A.row_ptr, A.col_ind, A.val, _i32(conn),
                              np.ascontiguousarray(aggregates, dtype=np.int64), relax,
                              C.byref(s))
        out = _struct_to_csr(s)
        self.L.or_csr_free(C.byref(s))
        return out

    def transpose(self, A):
        a = _struct_from(A)
        t = _CsrStruct()
        self.L.or_transpose(C.byref(a), C.byref(t))
        out = _struct_to_csr(t)
        self.L.or_csr_free(C.byref(t))
        return out

    def csrgemm(self, A, B):
        a, b = _struct_from(A), _struct_from(B)
        c = _CsrStruct()
        self.L.or_csrgemm(C.byref(a), C.byref(b), C.byref(c))
        out = _struct_to_csr(c)
        self.L.or_csr_free(C.byref(c))
        return out

    # ---- hierarchy / solve ----
    def saamg_setup(self, A, max_level=20, eps0=0.08, relax=2.0 / 3.0):
        h = self.L.or_saamg_setup(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, max_level,
                                  eps0, relax)
        return OracleHierarchy(self, h)

    def bicgstab(self, A, b, x0=None, precond=0, hier=None, rel_tol=1e-8, abs_tol=1e-300,
                 max_iter=1000):
        """bicgstab_solver::solve; precond 0 none, 1 Jacobi, 2 one SA-AMG V-cycle
        of hier (an OracleHierarchy); returns (x, iterations)."""
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        it = self.L.or_bicgstab(A.m, A.row_ptr, A.col_ind, A.val, x, _f64(b), precond,
                                hier.h if hier else None, rel_tol, abs_tol, max_iter)
        return x, it

    def gmres(self, A, b, restart=30, x0=None, precond=0, hier=None, rel_tol=1e-8,
              abs_tol=1e-300, max_iter=1000):
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        it = self.L.or_gmres(A.m, A.row_ptr, A.col_ind, A.val, x, _f64(b), restart, precond,
                             hier.h if hier else None, rel_tol, abs_tol, max_iter)
        return x, it

    def uaamg_setup(self, A, max_level=20, eps0=0.001):
        h = self.L.or_uaamg_setup(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, max_level, eps0)
        return OracleHierarchy(self, h)

    def jacobi_iteration(self, A, xold, b):
        x = np.zeros(A.m)
        self.L.or_jacobi_iteration(A.m, A.row_ptr, A.col_ind, A.val, x, _f64(xold), _f64(b))
        return x

    def cholesky_factor(self, dense):
        """Lower-triangular L with L L^T = dense (or_cholesky_factor)."""
        a = np.ascontiguousarray(dense, dtype=np.float64).copy()
        self.L.or_cholesky_factor(a.shape[0], a)
        return np.tril(a)

    def cg(self, A, b, x0=None, rel_tol=1e-6, abs_tol=1e-6, max_iter=1000):
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        rn = C.c_double(0.0)
        it = self.L.or_cg(A.m, A.row_ptr, A.col_ind, A.val, x, _f64(b), rel_tol, abs_tol,
                          max_iter, C.byref(rn))
        return x, it, rn.value


class OracleHierarchy:
    def __init__(self, o, h):
        self.o = o
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.o.L.or_hierarchy_free(self.h)
            self.h = None

    @property
    def levels(self):
        return self.o.L.or_hierarchy_levels(self.h)

    def get(self, level, kind):
        """kind: 'A', 'P' or 'R'."""
        k = {"A": 0, "P": 1, "R": 2}[kind]
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        if self.o.L.or_hierarchy_dims(self.h, level, k, C.byref(m), C.byref(n),
                                      C.byref(nnz)) != 0:
            return None
        rp = np.zeros(m.value + 1, dtype=np.int32)
        ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        v = np.zeros(max(nnz.value, 1))
        self.o.L.or_hierarchy_copy(self.h, level, k, rp, ci, v)
        return CSR(rp, ci[: nnz.value], v[: nnz.value], m.value, n.value)

    def precond(self, r, n1=1, n2=1, coarse_solver=0):
        z = np.zeros(len(r))
        self.o.L.or_amg_precond(self.h, _f64(r), z, n1, n2, coarse_solver)
        return z

    def amg_solve(self, b, x0=None, n1=1, n2=1, cycle=0, smoother=0, rel_tol=1e-8,
                  abs_tol=1e-300, max_cycle=100):
        """amg_solve: cycle 0 V / 1 W / 2 F, smoother 0 Jacobi / 1 GS / 2 SGS / 3 SOR /
        4 SSOR; returns (x, cycles, last residual inf-norm)."""
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        rn = C.c_double(0.0)
        k = self.o.L.or_amg_solve(self.h, x, _f64(b), n1, n2, cycle, smoother, rel_tol,
                                  abs_tol, max_cycle, C.byref(rn))
        return x, k, rn.value

    def pcg(self, b, x0=None, rel_tol=1e-8, abs_tol=1e-300, max_iter=1000, n1=1, n2=1,
            coarse_solver=0):
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        rn = C.c_double(0.0)
        it = self.o.L.or_pcg_saamg(self.h, x, _f64(b), rel_tol, abs_tol, max_iter, n1, n2,
                                   coarse_solver, C.byref(rn))
        return x, it, rn.value

    def pcg_op(self, A, b, x0=None, rel_tol=1e-8, abs_tol=1e-300, max_iter=1000, n1=1, n2=1,
               coarse_solver=0):
        """PCG on operator A (a CSR of the finest level's size) preconditioned
        by one V-cycle of this hierarchy."""
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        rn = C.c_double(0.0)
        it = self.o.L.or_pcg_saamg_op(A.m, A.row_ptr, A.col_ind, A.val, self.h, x, _f64(b),
                                      rel_tol, abs_tol, max_iter, n1, n2, coarse_solver,
                                      C.byref(rn))
        return x, it, rn.value


class Reference:
    """LinAlgLib's own host library, compiled from its sources by oracle/Makefile."""

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (make -C oracle ref)")
        L = C.CDLL(path)
        self.L = L
        L.ref_set_threads.argtypes = [C.c_int]
        L.ref_saamg_setup.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int]
        L.ref_saamg_setup.restype = C.c_void_p
        L.ref_hierarchy_free.argtypes = [C.c_void_p]
        L.ref_hierarchy_levels.argtypes = [C.c_void_p]
        L.ref_hierarchy_dims.argtypes = [C.c_void_p, C.c_int, C.c_int, _ip, _ip, _ip]
        L.ref_hierarchy_copy.argtypes = [C.c_void_p, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_amg_precond.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.c_int]
        L.ref_pcg_saamg_op.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_void_p,
                                       _f64p, _f64p, C.c_double, C.c_double, C.c_int, C.c_int,
                                       C.c_int]
        L.ref_pcg_saamg.argtypes = [C.c_void_p, _f64p, _f64p, C.c_double, C.c_double, C.c_int,
                                    C.c_int, C.c_int]
        L.ref_cg.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_double,
                             C.c_double, C.c_int]
        L.ref_csrmv.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, _i32p, _i32p, _f64p,
                                _f64p, C.c_double, _f64p]
        L.ref_residual.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p,
                                   _f64p, _f64p]
        L.ref_dot.argtypes = [C.c_int, _f64p, _f64p]
        L.ref_dot.restype = C.c_double
        L.ref_strength.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_double,
                                   _i32p]
        L.ref_aggregates.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, _i32p,
                                     _i64p]
        L.ref_csrgemm.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p] * 2
        L.ref_csrgemm.restype = C.c_void_p
        L.ref_transpose.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_transpose.restype = C.c_void_p
        L.ref_csr_dims.argtypes = [C.c_void_p, _ip, _ip, _ip]
        L.ref_csr_copy.argtypes = [C.c_void_p, _i32p, _i32p, _f64p]
        L.ref_csr_free.argtypes = [C.c_void_p]
        L.ref_uaamg_setup.argtypes = [C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p, C.c_int]
        L.ref_uaamg_setup.restype = C.c_void_p
        L.ref_amg_solve.argtypes = [C.c_void_p, _f64p, _f64p, C.c_int, C.c_int, C.c_int,
                                    C.c_int, C.c_double, C.c_double, C.c_int]
        L.ref_amg_solve.restype = C.c_int
        L.ref_bicgstab.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int,
                                   C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.ref_bicgstab.restype = C.c_int
        L.ref_gmres.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int,
                                C.c_int, C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.ref_gmres.restype = C.c_int
        L.ref_read_mtx.argtypes = [C.c_char_p]
        L.ref_read_mtx.restype = C.c_void_p
        L.ref_write_mtx.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _i32p, _i32p, _f64p]
        L.ref_write_mtx.restype = C.c_int

    @staticmethod
    def available():
        return os.path.exists(REF_SO)

    def set_threads(self, n):
        self.L.ref_set_threads(int(n))

    def _take(self, p):
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        self.L.ref_csr_dims(p, C.byref(m), C.byref(n), C.byref(nnz))
        rp = np.zeros(m.value + 1, dtype=np.int32)
        ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        v = np.zeros(max(nnz.value, 1))
        self.L.ref_csr_copy(p, rp, ci, v)
        self.L.ref_csr_free(p)
        return CSR(rp, ci[: nnz.value], v[: nnz.value], m.value, n.value)

    def csrmv(self, A, x, alpha=1.0, beta=0.0, y=None):
        y = np.zeros(A.m) if y is None else _f64(y).copy()
        self.L.ref_csrmv(A.m, A.n, A.nnz, alpha, A.row_ptr, A.col_ind, A.val, _f64(x), beta, y)
        return y

    def residual(self, A, x, b):
        r = np.zeros(A.m)
        self.L.ref_residual(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, _f64(x), _f64(b), r)
        return r

    def dot(self, x, y):
        return self.L.ref_dot(len(x), _f64(x), _f64(y))

    def strength(self, A, eps):
        conn = np.zeros(A.nnz, dtype=np.int32)
        self.L.ref_strength(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, eps, conn)
        return conn

    def aggregates(self, A, conn):
        agg = np.zeros(A.m, dtype=np.int64)
        self.L.ref_aggregates(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, _i32(conn), agg)
        return agg

    def csrgemm(self, A, B):
        p = self.L.ref_csrgemm(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, B.m, B.n, B.nnz,
                               B.row_ptr, B.col_ind, B.val)
        return self._take(p)

    def transpose(self, A):
        return self._take(self.L.ref_transpose(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val))

    def bicgstab(self, A, b, x0=None, precond=0, hier=None, rel_tol=1e-8, abs_tol=1e-300,
                 max_iter=1000):
        """bicgstab_solver::solve; precond 0 none, 1 Jacobi, 2 one SA-AMG V-cycle of
        hier (a RefHierarchy)."""
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        it = self.L.ref_bicgstab(A.m, A.nnz, A.row_ptr, A.col_ind, A.val, x, _f64(b), precond,
                                 hier.h if hier else None, rel_tol, abs_tol, max_iter)
        return x, it

    def gmres(self, A, b, restart=30, x0=None, precond=0, hier=None, rel_tol=1e-8,
              abs_tol=1e-300, max_iter=1000):
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        it = self.L.ref_gmres(A.m, A.nnz, A.row_ptr, A.col_ind, A.val, x, _f64(b), restart,
                              precond, hier.h if hier else None, rel_tol, abs_tol, max_iter)
        return x, it

    def read_mtx(self, path):
        """csr_matrix::read_mtx of the reference; None when it refuses the file."""
        p = self.L.ref_read_mtx(str(path).encode())
        return self._take(p) if p else None

    def write_mtx(self, path, A):
        ci = A.col_ind if A.nnz else np.zeros(1, np.int32)
        v = A.val if A.nnz else np.zeros(1)
        return bool(self.L.ref_write_mtx(str(path).encode(), A.m, A.n, A.nnz, A.row_ptr,
                                         _i32(ci), _f64(v)))

    def saamg_setup(self, A, max_level=20):
        h = self.L.ref_saamg_setup(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, max_level)
        return RefHierarchy(self, h)

    def uaamg_setup(self, A, max_level=20):
        h = self.L.ref_uaamg_setup(A.m, A.n, A.nnz, A.row_ptr, A.col_ind, A.val, max_level)
        return RefHierarchy(self, h)

    def cg(self, A, b, x0=None, rel_tol=1e-6, abs_tol=1e-6, max_iter=1000):
        x = np.zeros(A.m) if x0 is None else _f64(x0).copy()
        it = self.L.ref_cg(A.m, A.nnz, A.row_ptr, A.col_ind, A.val, x, _f64(b), rel_tol,
                           abs_tol, max_iter)
        return x, it


class RefHierarchy:
    def __init__(self, r, h):
        self.r = r
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            self.r.L.ref_hierarchy_free(self.h)
            self.h = None

    @property
    def levels(self):
        return self.r.L.ref_hierarchy_levels(self.h)

    def get(self, level, kind):
        k = {"A": 0, "P": 1, "R": 2}[kind]
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        if self.r.L.ref_hierarchy_dims(self.h, level, k, C.byref(m), C.byref(n),
                                       C.byref(nnz)) != 0:
            return None
        rp = np.zeros(m.value + 1, dtype=np.int32)
        ci = np.zeros(max(nnz.value, 1), dtype=np.int32)
        v = np.zeros(max(nnz.value, 1))
        self.r.L.ref_hierarchy_copy(self.h, level, k, rp, ci, v)
        return CSR(rp, ci[: nnz.value], v[: nnz.value], m.value, n.value)

    def precond(self, r, n1=1, n2=1):
        z = np.zeros(len(r))
        self.r.L.ref_amg_precond(self.h, _f64(r), z, n1, n2)
        return z

    def amg_solve(self, b, x0=None, n1=1, n2=1, cycle=0, smoother=0, rel_tol=1e-8,
                  abs_tol=1e-300, max_cycle=100):
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        k = self.r.L.ref_amg_solve(self.h, x, _f64(b), n1, n2, cycle, smoother, rel_tol,
                                   abs_tol, max_cycle)
        return x, k

    def pcg(self, b, x0=None, rel_tol=1e-8, abs_tol=1e-300, max_iter=1000, n1=1, n2=1):
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        it = self.r.L.ref_pcg_saamg(self.h, x, _f64(b), rel_tol, abs_tol, max_iter, n1, n2)
        return x, it

    def pcg_op(self, A, b, x0=None, rel_tol=1e-8, abs_tol=1e-300, max_iter=1000, n1=1, n2=1):
        x = np.zeros(len(b)) if x0 is None else _f64(x0).copy()
        it = self.r.L.ref_pcg_saamg_op(A.m, A.nnz, A.row_ptr, A.col_ind, A.val, self.h, x,
                                       _f64(b), rel_tol, abs_tol, max_iter, n1, n2)
        return x, it
