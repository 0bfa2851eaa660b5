/*
 * CPU oracle for the SA-AMG-PCG hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of LinAlgLib's host algorithms (file:line
 * citations on every function in lal_oracle.c).  It exists to check the HIP
 * library: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path never links or calls it.
 *
 * Arithmetic follows the reference exactly: compiled with -ffp-contract=off
 * and no -march, explicit std::fma is restated as fma(), every other product
 * and sum is rounded separately, and every loop keeps the reference's order.
 */
#ifndef LAL_ORACLE_H
#define LAL_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
    int     m, n, nnz;
    int*    row_ptr; /* m+1 */
    int*    col_ind; /* nnz */
    double* val;     /* nnz */
} or_csr;

void or_csr_free(or_csr* A);
void or_set_threads(int n);

/* ---- BLAS-1 / SpMV ---- */
void   or_csrmv(int m, int n, int nnz, double alpha, const int* rp, const int* ci,
                const double* v, const double* x, double beta, double* y);
void   or_residual(int m, const int* rp, const int* ci, const double* v, const double* x,
                   const double* b, double* r);
double or_dot(int n, const double* x, const double* y);
double or_norm_inf(int n, const double* x);
void   or_axpy(int n, double alpha, const double* x, double* y);
void   or_axpby(int n, double alpha, const double* x, double beta, double* y);
void   or_diagonal(int m, const int* rp, const int* ci, const double* v, double* d);
void   or_exclusive_scan_i64(int n, int64_t* x);
void   or_exclusive_scan_i32(int n, int32_t* x);

/* ---- SA-AMG setup pieces ---- */
void or_strength(int m, const int* rp, const int* ci, const double* v, double eps, int* conn);
unsigned int or_hash1(unsigned int x);
void or_pmis_init(int m, const int* rp, const int* conn, int* state, int* hash);
void or_pmis_round(int m, const int* rp, const int* ci, const int* conn, const int* state,
                   const int* hash, int64_t* aggregates, int* max_state, int* complete);
void or_pmis_join(int m, const int* rp, const int* ci, const int* conn, const int* state,
                  int64_t* aggregates, int* max_state);
void or_aggregates_pmis(int m, const int* rp, const int* ci, const int* conn,
                        int64_t* aggregates, int64_t* root_nodes);
/* P is allocated by the oracle; free with or_csr_free. */
void or_prolongator(int m, const int* rp, const int* ci, const double* v, const int* conn,
                    const int64_t* aggregates, double relax, or_csr* P);
void or_transpose(const or_csr* A, or_csr* T);
void or_csrgemm(const or_csr* A, const or_csr* B, or_csr* C);

/* ---- hierarchy ---- */
#define OR_MAX_LEVELS 20
typedef struct
{
    or_csr A[OR_MAX_LEVELS + 1];
    or_csr P[OR_MAX_LEVELS];
    or_csr R[OR_MAX_LEVELS];
    int    total_levels;
} or_hierarchy;

or_hierarchy* or_saamg_setup(int m, int n, int nnz, const int* rp, const int* ci,
                             const double* v, int max_level, double eps0, double relax);
/* uaamg_setup (uaamg.cpp:125-224): unsmoothed prolongator, eps from eps0 (0.001 there). */
or_hierarchy* or_uaamg_setup(int m, int n, int nnz, const int* rp, const int* ci,
                             const double* v, int max_level, double eps0);
void or_ua_prolongator(int m, const int64_t* agg, or_csr* P);
void          or_hierarchy_free(or_hierarchy* h);
int           or_hierarchy_levels(const or_hierarchy* h);
/* kind: 0 = A, 1 = P, 2 = R.  Returns -1 when the level does not exist. */
int  or_hierarchy_dims(const or_hierarchy* h, int level, int kind, int* m, int* n, int* nnz);
void or_hierarchy_copy(const or_hierarchy* h, int level, int kind, int* rp, int* ci, double* v);

/* ---- solve ---- */
void or_jacobi_iteration(int m, const int* rp, const int* ci, const double* v, double* x,
                         const double* xold, const double* b);
/* Dense Cholesky factor, in place on a row-major N x N SPD matrix (lower
 * triangle = L with L L^T = A). */
void or_cholesky_factor(int N, double* L);
/* coarse_solver: 0 = 1000 Gauss-Seidel sweeps (reference), 1 = dense Cholesky */
void or_vcycle(const or_hierarchy* h, double* x, const double* b, int n1, int n2, int level,
               int coarse_solver);
/* z = M^-1 r : z zeroed, one V-cycle with undamped Jacobi smoothing. */
/* One smoothing step on A x = b (amg.cpp:110-142): smoother 0 Jacobi, 1 Gauss-Seidel,
 * 2 "symmetric" Gauss-Seidel (a forward Gauss-Seidel sweep, as written), 3 SOR and
 * 4 SSOR with omega = 0.5. */
void or_smooth(int m, const int* rp, const int* ci, const double* v, double* x,
               const double* b, int smoother);
/* One V (0), W (1) or F (2) cycle on level `level` (amg.cpp:144-513), x in place;
 * n3 is the W/F cycles' post-smoothing count. */
void or_cycle(const or_hierarchy* h, double* x, const double* b, int n1, int n2, int n3,
              int level, int cycle, int smoother);
/* amg_solve (amg.cpp:516-571): cycles until the inf-norm residual converges or more
 * than max_cycle cycles ran; returns the cycle count, *res_norm the last norm. */
int or_amg_solve(const or_hierarchy* h, double* x, const double* b, int n1, int n2, int cycle,
                 int smoother, double rel_tol, double abs_tol, int max_cycle,
                 double* res_norm);
/* bicgstab_solver (krylov/bicgstab.cpp:64-231); kind 0 none, 1 Jacobi, 2 one SA-AMG
 * V-cycle of h (1 + 1 sweeps).  Returns the iteration count, x in place. */
int or_bicgstab(int m, const int* rp, const int* ci, const double* v, double* x,
                const double* b, int kind, const or_hierarchy* h, double rel_tol,
                double abs_tol, int max_iter);
/* Restarted left-preconditioned GMRES in the form of krylov/gmres.cpp (see the .c). */
int or_gmres(int m, const int* rp, const int* ci, const double* v, double* x, const double* b,
             int restart, int kind, const or_hierarchy* h, double rel_tol, double abs_tol,
             int max_iter);
void or_amg_precond(const or_hierarchy* h, const double* r, double* z, int n1, int n2,
                    int coarse_solver);
/* Returns the iteration count as cg_solver::solve_precond does; res_norm gets the
 * final recursively-updated infinity norm. */
int or_pcg_saamg(const or_hierarchy* h, double* x, const double* b, double rel_tol,
                 double abs_tol, int max_iter, int n1, int n2, int coarse_solver,
                 double* res_norm);
/* The same loop with the Krylov products on (m, rp, ci, v) and the V-cycle
 * of h (built for any matrix of that size). */
int or_pcg_saamg_op(int m, const int* rp, const int* ci, const double* v, const or_hierarchy* h,
                    double* x, const double* b, double rel_tol, double abs_tol, int max_iter,
                    int n1, int n2, int coarse_solver, double* res_norm);
int or_cg(int m, const int* rp, const int* ci, const double* v, double* x, const double* b,
          double rel_tol, double abs_tol, int max_iter, double* res_norm);

#ifdef __cplusplus
}
#endif
#endif
