/*
 * lal.h -- C ABI of the MI355X (gfx950) backend for LinAlgLib's SA-AMG-PCG path.
 *
 * Every entry point takes raw device pointers and int32 sizes, returns an int
 * status (LAL_OK = 0) and is asynchronous on the handle's stream unless it
 * hands a scalar back to the host.  The caller owns every buffer it passes;
 * descriptors and hierarchies are opaque handles with create/destroy pairs.
 *
 * Each declaration names the reference interface it replaces.  Paths are
 * relative to LinAlgLib's library/src/ tree.  In the reference those device
 * entry points are `cuda_*` launchers called from the `device_*` shims under
 * backend/device/; a maintainer swaps each shim body for the lal_* call.
 */
#ifndef LAL_H
#define LAL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* status                                                                   */
/* ------------------------------------------------------------------------ */
enum
{
    LAL_OK                = 0,
    LAL_ERR_INVALID       = 1, /* bad argument (null pointer, negative size, ...) */
    LAL_ERR_OOM           = 2, /* device allocation failed                        */
    LAL_ERR_HIP           = 3, /* a HIP runtime call or kernel launch failed      */
    LAL_ERR_NOT_CONVERGED = 4, /* iteration limit reached                         */
    LAL_ERR_NAN           = 5, /* NaN/Inf met in a residual norm                  */
    LAL_ERR_NO_DEVICE     = 6, /* no usable gfx950 device                         */
    LAL_ERR_UNSUPPORTED   = 7  /* input outside what the backend supports         */
};

typedef struct lal_context* lal_handle;

/* Human-readable text for a status code. */
const char* lal_status_string(int status);
/* Detail of the last failure on the calling thread (the reference prints
 * its errors instead, common.cuh:58-97). */
const char* lal_last_error(void);

/* ------------------------------------------------------------------------ */
/* context and memory -- replaces backend/device/device_memory.h:34-52        */
/* (device_allocate, device_free, copy_h2d, copy_d2h, copy_d2d, device_fill,  */
/* device_sync) and linalg_memory.cpp:37-45 (synchronize).                    */
/* ------------------------------------------------------------------------ */

/* Creates a context on the current device with its own stream (default
 * priority), pinned scalar read-back memory and a small reduction workspace. */
int lal_create(lal_handle* handle);
int lal_destroy(lal_handle handle);
/* Use a caller-owned hipStream_t (passed as void*) instead of the own one. */
int lal_set_stream(lal_handle handle, void* stream);
int lal_get_stream(lal_handle handle, void** stream);
int lal_sync(lal_handle handle);

int lal_alloc(void** ptr, size_t bytes);
int lal_free(void* ptr);
int lal_memcpy_h2d(lal_handle handle, void* dst, const void* src, size_t bytes);
int lal_memcpy_d2h(lal_handle handle, void* dst, const void* src, size_t bytes);
int lal_memcpy_d2d(lal_handle handle, void* dst, const void* src, size_t bytes);
int lal_fill_f64(lal_handle handle, double* x, size_t n, double value);
int lal_fill_f32(lal_handle handle, float* x, size_t n, float value);
int lal_fill_i32(lal_handle handle, int32_t* x, size_t n, int32_t value);
int lal_fill_i64(lal_handle handle, int64_t* x, size_t n, int64_t value);
/* Elementwise precision conversion (FP32 copies of FP64 operators/vectors). */
int lal_convert_f64_f32(lal_handle handle, const double* x, float* y, size_t n);
int lal_convert_f32_f64(lal_handle handle, const float* x, double* y, size_t n);

/* ------------------------------------------------------------------------ */
/* BLAS-1 -- replaces backend/device/device_axpy.h:33-41 and                 */
/* device_math.h:54-60 (cuda_axpy / cuda_axpby / cuda_axpbypgz /             */
/* cuda_dot_product / cuda_norm_inf) and device_primitives.h:38-47           */
/* (cuda_exclusive_scan).  Same arithmetic as backend/host/host_axpy.cpp and  */
/* host_math.cpp.  Reductions are deterministic (fixed order).               */
/* ------------------------------------------------------------------------ */
int lal_axpy_f64(lal_handle h, int n, double alpha, const double* x, double* y);
int lal_axpby_f64(lal_handle h, int n, double alpha, const double* x, double beta, double* y);
int lal_axpbypgz_f64(lal_handle h, int n, double alpha, const double* x, double beta,
                     const double* y, double gamma, double* z);
int lal_axpy_f32(lal_handle h, int n, float alpha, const float* x, float* y);
int lal_axpby_f32(lal_handle h, int n, float alpha, const float* x, float beta, float* y);
/* Result returned to host memory (blocks until ready). FP64 accumulation.
 * lal_dot_f64 with n <= 1024 sums in index order (bit-identical to the host
 * loop at one OpenMP thread); longer vectors use a fixed-order tree. */
int lal_dot_f64(lal_handle h, int n, const double* x, const double* y, double* result);
int lal_dot_f32(lal_handle h, int n, const float* x, const float* y, double* result);
/* Fused y = alpha x + y (lal_axpy_f64's arithmetic) and *result = <y, z>
 * over the updated y (z may be y), in one pass: the CG residual update and
 * its inner product (cg.cpp:177-203).  Bit-identical to lal_axpy_f64 followed
 * by lal_dot_f64. */
int lal_axpy_dot_f64(lal_handle h, int n, double alpha, const double* x, double* y,
                     const double* z, double* result);
int lal_norm_inf_f64(lal_handle h, int n, const double* x, double* result);
int lal_norm_inf_f32(lal_handle h, int n, const float* x, double* result);
/* In-place exclusive prefix sum. */
int lal_exclusive_scan_i32(lal_handle h, int n, int32_t* x);
int lal_exclusive_scan_i64(lal_handle h, int n, int64_t* x);
/* lal_exclusive_scan_i64 without any allocation.  No counterpart in
 * LinAlgLib: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Same kernels and result as lal_exclusive_scan_i64, bit for bit; the tile
 * sums live in a static per-device array of LAL_SCAN_NOALLOC_MAX_N / 2048
 * int64 entries (256 KiB), so n > LAL_SCAN_NOALLOC_MAX_N is refused with
 * LAL_ERR_INVALID.  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams or two handles at once) race. */
#define LAL_SCAN_NOALLOC_MAX_N 67108864
int lal_exclusive_scan_i64_noalloc(lal_handle h, int n, int64_t* x);

/* ------------------------------------------------------------------------ */
/* CSR SpMV -- replaces backend/device/device_matrix_vector.h:35-45          */
/* (cuda_csrmv_analysis, cuda_csrmv_solve, cuda_compute_residual;            */
/* backend/device/cuda/cuda_matrix_vector.h:34-80).  y = alpha*A*x + beta*y   */
/* with beta honoured (the CUDA path ignores it).  Row-length-adaptive:      */
/* the analysis packs rows into LDS-staged blocks.  With exact = 1 every row */
/* is summed by one lane as s = fma(val, x, s) in CSR order, which is        */
/* bit-identical to backend/host/host_matrix_vector.cpp:59-84; with exact = 0 */
/* rows in sparse blocks are split over a lane group (deterministic, within  */
/* 1e-14 relative of the sequential sum).                                    */
/* ------------------------------------------------------------------------ */
typedef struct lal_csrmv_info* lal_csrmv_descr;

/* Validates the pattern (row_ptr monotone from 0 to nnz, every column in
 * [0, n)) and packs the rows into blocks; refuses an invalid matrix with
 * LAL_ERR_INVALID before any kernel could read outside x. */
int lal_csrmv_analysis(lal_handle h, int m, int n, int nnz, const int* row_ptr,
                       const int* col_ind, int exact, lal_csrmv_descr* descr);
int lal_csrmv_descr_destroy(lal_csrmv_descr descr);
/* descr may be NULL: rows are then taken 256 per block and each is summed by
 * one lane in CSR order, as with an exact descriptor. */
int lal_csrmv_f64(lal_handle h, lal_csrmv_descr descr, int m, int n, int nnz, double alpha,
                  const int* row_ptr, const int* col_ind, const double* val, const double* x,
                  double beta, double* y);
int lal_csrmv_f32(lal_handle h, lal_csrmv_descr descr, int m, int n, int nnz, float alpha,
                  const int* row_ptr, const int* col_ind, const float* val, const float* x,
                  float beta, float* y);
/* r = b - A x with s += val*x (two roundings), as host_matrix_vector.cpp:35-57. */
int lal_residual_f64(lal_handle h, lal_csrmv_descr descr, int m, int n, int nnz,
                     const int* row_ptr, const int* col_ind, const double* val,
                     const double* x, const double* b, double* r);
/* The same in FP32 (values, x, b, r), for the FP32 copies of an operator. */
int lal_residual_f32(lal_handle h, lal_csrmv_descr descr, int m, int n, int nnz,
                     const int* row_ptr, const int* col_ind, const float* val, const float* x,
                     const float* b, float* r);
/* d_i = first a_ii of row i; untouched when absent (host_extract.cpp:34-52). */
int lal_diagonal_f64(lal_handle h, int m, const int* row_ptr, const int* col_ind,
                     const double* val, double* d);

/* ------------------------------------------------------------------------ */
/* Sparse matrix products -- replaces backend/device/device_csr2csc.h:33     */
/* (transpose) and device_csrgemm.h:37-44 (csrgemm_nnz / csrgemm_solve).     */
/* ------------------------------------------------------------------------ */
/* T = A^T (A is m x n; T is n x m with nnz entries; caller allocates T's
 * n+1 / nnz / nnz arrays).  Columns within each row of T ascend
 * (host_csr2csc.cpp:30-75). */
int lal_transpose_f64(lal_handle h, int m, int n, int nnz, const int* row_ptr,
                      const int* col_ind, const double* val, int* t_row_ptr, int* t_col_ind,
                      double* t_val);
/* The same with FP32 values (the pattern is that of lal_transpose_f64). */
int lal_transpose_f32(lal_handle h, int m, int n, int nnz, const int* row_ptr,
                      const int* col_ind, const float* val, int* t_row_ptr, int* t_col_ind,
                      float* t_val);
/* C = A B, two passes.  Pass 1 fills row_ptr_C (m+1, device) and returns nnz
 * to *nnz_C (host).  Pass 2 fills col_ind_C / val_C with columns ascending
 * and every entry folded in A-row-then-B-row order from its first product,
 * bit-identical to backend/host/host_csrgemm.cpp:91-183.  Both passes check
 * the operands first: a column index of A outside [0, k) or of B outside
 * [0, n) gives LAL_ERR_INVALID, and a row of B that repeats a column index
 * gives LAL_ERR_UNSUPPORTED (the host loop would sum the repeats). */
int lal_csrgemm_nnz(lal_handle h, int m, int k, int n, int nnz_A, const int* row_ptr_A,
                    const int* col_ind_A, int nnz_B, const int* row_ptr_B,
                    const int* col_ind_B, int* row_ptr_C, int* nnz_C);
int lal_csrgemm_solve_f64(lal_handle h, int m, int k, int n, int nnz_A, const int* row_ptr_A,
                          const int* col_ind_A, const double* val_A, int nnz_B,
                          const int* row_ptr_B, const int* col_ind_B, const double* val_B,
                          const int* row_ptr_C, int* col_ind_C, double* val_C);

/* ------------------------------------------------------------------------ */
/* SA-AMG setup pieces -- replaces backend/device/device_amg_strength.h:35   */
/* (an empty stub on CUDA) and device_amg_aggregation.h:40-53, with the host */
/* semantics of backend/host/host_amg_strength.cpp:33-70 and                 */
/* host_amg_aggregation.cpp:32-221.                                          */
/* ------------------------------------------------------------------------ */
int lal_sa_strength_f64(lal_handle h, int m, int nnz, const int* row_ptr, const int* col_ind,
                        const double* val, double eps, int* connections);
/* The whole of iterative_solvers/amg/amg_aggregation.cpp:106-215
 * (compute_aggregates_using_pmis): PMIS rounds (at most 20), root flags,
 * exclusive scan of the aggregate flags, two join passes.  `aggregates` is
 * written completely (zeroing is done here). */
int lal_pmis_aggregate(lal_handle h, int m, int nnz, const int* row_ptr, const int* col_ind,
                       const int* connections, int64_t* aggregates);
/* The three PMIS steps one to one, for a driver that runs
 * amg_aggregation.cpp:106-215 itself:
 *   lal_pmis_init  -- device_initialize_pmis_state (device_amg_aggregation.h:40-43;
 *                     host_amg_aggregation.cpp:32-68): state_i = 0 when row i
 *                     has a strong connection, else -2; hash_i = hash1(i).
 *   lal_pmis_round -- device_find_maximum_distance_two_node
 *                     (device_amg_aggregation.h:45-51; host_amg_aggregation.cpp:
 *                     77-180): one round on the snapshot `state`; an
 *                     undecided row that is the lexicographic (state, hash)
 *                     maximum of its distance-two strong neighbourhood becomes
 *                     a root (aggregates 1, max_state 1), one whose maximum is
 *                     a root becomes -1 (aggregates 0); *complete (host
 *                     memory, the call blocks until it is known) is 1 when no
 *                     row was left undecided, as the reference's bool.
 *   lal_pmis_join  -- device_add_unassigned_nodes_to_closest_aggregation
 *                     (device_amg_aggregation.h:53-58; host_amg_aggregation.cpp:
 *                     182-221): a state -1 row takes the aggregate of its first
 *                     strong state-1 neighbour in CSR order and max_state 1; a
 *                     state -2 row gets aggregate -2.  (The reference's
 *                     aggregate_root_nodes argument is unused there.)
 * The driver copies max_state into state before each round and each join,
 * and scans the aggregate flags in between (lal_exclusive_scan_i64). */
int lal_pmis_init(lal_handle h, int m, int nnz, const int* row_ptr, const int* connections,
                  int* state, int* hash);
int lal_pmis_round(lal_handle h, int m, int nnz, const int* row_ptr, const int* col_ind,
                   const int* connections, const int* state, const int* hash,
                   int64_t* aggregates, int* max_state, int* complete);
int lal_pmis_join(lal_handle h, int m, int nnz, const int* row_ptr, const int* col_ind,
                  const int* connections, const int* state, int64_t* aggregates,
                  int* max_state);
/* P of iterative_solvers/amg/saamg.cpp:51-207 in two passes: the first
 * fills p_row_ptr (m+1) and returns nnz_P and the column count n_c; the
 * second fills p_col_ind / p_val. */
int lal_sa_prolongator_nnz(lal_handle h, int m, int nnz, const int* row_ptr,
                           const int* col_ind, const int* connections,
                           const int64_t* aggregates, int* p_row_ptr, int* nnz_P, int* n_c);
int lal_sa_prolongator_fill_f64(lal_handle h, int m, int nnz, const int* row_ptr,
                                const int* col_ind, const double* val, const int* connections,
                                const int64_t* aggregates, double relax, const int* p_row_ptr,
                                int* p_col_ind, double* p_val);

/* ------------------------------------------------------------------------ */
/* Solver level: hierarchy, preconditioner, PCG -- the device counterparts   */
/* of iterative_solvers/amg/saamg.cpp:210-282 (saamg_setup),                 */
/* amg.cpp:144-267 (vcycle), krylov/cg.cpp:136-223 (solve_precond) and       */
/* iter_control.cpp:38-91.  The reference runs these on the host only.       */
/* ------------------------------------------------------------------------ */
typedef struct lal_hierarchy_t* lal_hierarchy;

typedef struct
{
    int    max_levels;     /* capped at 20 as saamg.cpp:216                  */
    double eps;            /* strength threshold at level 0 (0.08)           */
    double relax;          /* prolongator smoothing weight (2/3)             */
    int    coarse_solver;  /* 0: 1000 Gauss-Seidel sweeps as amg.cpp:238-264;
                              1: dense Cholesky factored at setup            */
    int    exact;          /* 1: every SpMV sums rows sequentially (bitwise
                              equal to the host path)                         */
    int    build_f32;      /* 1: also keep FP32 copies of every level        */
} lal_saamg_options;

void lal_saamg_default_options(lal_saamg_options* opts);
/* uaamg_setup (iterative_solvers/amg/uaamg.cpp:125-224): the same level loop
 * with the unsmoothed prolongator (one 1.0 per aggregated row, column =
 * its aggregate; uaamg.cpp:48-123) and eps starting at 0.001 (the defaults
 * of lal_uaamg_default_options; `relax` is unused).  The hierarchy serves
 * lal_amg_solve, lal_amg_precond_f64 and the PCG entry points like a
 * smoothed one. */
void lal_uaamg_default_options(lal_saamg_options* opts);
int  lal_uaamg_setup(lal_handle h, int m, int n, int nnz, const int* row_ptr, const int* col_ind,
                     const double* val, const lal_saamg_options* opts, lal_hierarchy* hier);

/* Builds the hierarchy on the device from a device-resident CSR matrix. The
 * hierarchy keeps its own copy of A. */
int lal_saamg_setup(lal_handle h, int m, int n, int nnz, const int* row_ptr,
                    const int* col_ind, const double* val, const lal_saamg_options* opts,
                    lal_hierarchy* out);
/* May also be called after lal_destroy of the context it was built on (the
 * hierarchy is then only freed); no other call takes such a hierarchy. */
int lal_hierarchy_destroy(lal_hierarchy hier);
int lal_hierarchy_levels(lal_hierarchy hier, int* total_levels);
/* kind: 0 = A, 1 = P, 2 = R.  Device pointers stay owned by the hierarchy. */
int lal_hierarchy_level(lal_hierarchy hier, int level, int kind, int* m, int* n, int* nnz,
                        const int** row_ptr, const int** col_ind, const double** val);
/* Bytes per column index that the fast-path kernels read for A (kind 0), P
 * (1) or R (2) at a level: 2 when the setup also stored that operator's
 * columns as 16-bit offsets from each row block's smallest column (possible
 * when no block spans more than 65535 columns), else 4.  col_ind itself is
 * always int32.  No counterpart in LinAlgLib: it describes this backend's
 * storage, for byte accounting (bench.py's roofline of a solve). */
int lal_hierarchy_column_bytes(lal_hierarchy hier, int level, int kind, int* bytes);

/* z = M^-1 r: z zeroed, one V-cycle with n1 pre and n2 post undamped Jacobi
 * sweeps -- the reference composition
 * amg_solve(h, z, r, n1, n2, vcycle, jacobi, {max_cycle = 0}). */
int lal_amg_precond_f64(lal_handle h, lal_hierarchy hier, const double* r, double* z, int n1,
                        int n2);

typedef struct
{
    double rel_tol;   /* iter_control::rel_tol  */
    double abs_tol;   /* iter_control::abs_tol  */
    int    max_iter;  /* iter_control::max_iter */
    int    n1, n2;    /* pre/post Jacobi sweeps per level */
    /* mixed precision (lal_ir_solve): inner FP32 PCG tolerance and limits */
    double inner_rel_tol;
    int    inner_max_iter;
    int    max_outer;
    /* smoother of the FP64 PCG V-cycle (lal_pcg_solve_f64 / _op_f64 and the
       FP64 adjoint): 0 undamped Jacobi (the reference composition; default),
       or any lal_amg_solve smoother 1-6, e.g. 6 for symmetric multicolour
       Gauss-Seidel.  The mixed-precision schemes take 0 only. */
    int    smoother;
} lal_solve_options;

typedef struct
{
    int    iterations;       /* PCG: as cg_solver::solve_precond returns it;
                                IR: total inner iterations                   */
    int    outer_iterations; /* IR only                                     */
    int    converged;
    double initial_res_norm; /* inf-norm of b - A x0                         */
    double final_res_norm;   /* recursively updated inf-norm (PCG), true
                                inf-norm (IR)                                */
} lal_solve_info;

void lal_solve_default_options(lal_solve_options* opts);

/* SA-AMG-preconditioned CG in FP64: x holds the initial guess on entry. */
int lal_pcg_solve_f64(lal_handle h, lal_hierarchy hier, const double* b, double* x,
                      const lal_solve_options* opts, lal_solve_info* info);
/* cg_solver::solve(A, x, b, precond, control) (krylov/cg.cpp:136-223) with an
 * SA-AMG preconditioner as the reference takes them, independently: the
 * Krylov products q = A p and the initial residual use the caller's square
 * FP64 CSR matrix A (descr from lal_csrmv_analysis, or NULL for one lane per
 * row), the preconditioner is one V-cycle of hier, which may have been built
 * for another matrix of the same size.  lal_pcg_solve_f64 is this call with
 * A = the matrix hier was built from (its own resident copy). */
int lal_pcg_solve_op_f64(lal_handle h, lal_csrmv_descr descr, int m, int nnz,
                         const int* row_ptr, const int* col_ind, const double* val,
                         lal_hierarchy hier, const double* b, double* x,
                         const lal_solve_options* opts, lal_solve_info* info);
/* Mixed precision: FP64 outer residual r = b - A x and update x += d, FP32
 * inner SA-AMG-PCG (FP64 inner products) to inner_rel_tol; needs build_f32. */
int lal_ir_solve(lal_handle h, lal_hierarchy hier, const double* b, double* x,
                 const lal_solve_options* opts, lal_solve_info* info);
/* Mixed precision: cg_solver::solve_precond (iterative_solvers/krylov/cg.cpp:
 * 136-223) with x, r, p, A p and all inner products in FP64 and the SA-AMG
 * V-cycle (amg.cpp:144-267) applied on the FP32 levels; needs build_f32.
 * Same options and convergence test as lal_pcg_solve_f64; residual entries
 * must stay within FP32's normal range. */
int lal_pcg_solve_mixed(lal_handle h, lal_hierarchy hier, const double* b, double* x,
                        const lal_solve_options* opts, lal_solve_info* info);
/* Mixed precision: cg_solver::solve_precond (cg.cpp:136-223) with the Krylov
 * recurrences in FP32 (A's FP32 values, FP32 r/p/q/z and correction y, FP64
 * inner-product accumulation, the FP32 V-cycle) and reliable updates in FP64:
 * whenever the recursive residual has fallen by delta = opts->inner_rel_tol
 * (1e-2 when opts is NULL; the default options' 1e-4 suits well-conditioned
 * operators only) since the last update, or meets the tolerance, x += y in
 * FP64 and the residual is
 * replaced by the true FP64 b - A x.  Convergence (rel_tol / abs_tol on the
 * inf-norm) is tested on that true residual only; info->iterations counts CG
 * iterations as cg_solver does, info->outer_iterations the updates.  Needs
 * build_f32.  No reference counterpart beyond cg.cpp: the reference has no
 * mixed-precision solver. */
int lal_pcg_solve_reliable(lal_handle h, lal_hierarchy hier, const double* b, double* x,
                           const lal_solve_options* opts, lal_solve_info* info);

/* AMG as a stand-alone solver -- replaces linalg::amg_solve
 * (iterative_solvers/amg/amg.cpp:516-571; enums of include/iterative_solvers/
 * amg/amg.h:155-240).  cycle: 0 V, 1 W, 2 F; smoother: 0 Jacobi,
 * 1 Gauss-Seidel, 2 symmetric Gauss-Seidel (a forward sweep, as the reference
 * runs it), 3 SOR, 4 SSOR (omega = 0.5); and, with no reference counterpart,
 * 5 multicolour Gauss-Seidel (rows greedily coloured so that no two coupled
 * rows share a colour, colours swept in order, each colour one parallel pass)
 * and 6 symmetric multicolour Gauss-Seidel (forward then backward colour
 * order; a V-cycle with it is a symmetric preconditioner).  The colouring is
 * built on first use per level.  x is the initial guess on entry;
 * cycles run until the inf-norm residual meets rel_tol / abs_tol or more
 * than opts->max_iter cycles ran (iter_control::max_cycle);
 * info->iterations is the reference's return value.  Gauss-Seidel-type sweeps
 * run as level sets of the rows' dependency graph (built on first use per
 * level), bit-identical to the sequential sweep; with a hierarchy built in
 * exact mode the whole solve is bit-identical to the reference's. */
int lal_amg_solve(lal_handle h, lal_hierarchy hier, const double* b, double* x, int n1, int n2,
                  int cycle, int smoother, const lal_solve_options* opts, lal_solve_info* info);

/* ------------------------------------------------------------------------ */
/* Jacobi preconditioner and the nonsymmetric Krylov solvers -- replace      */
/* linalg::jacobi_precond (iterative_solvers/preconditioner/                 */
/* preconditioner.cpp:41-58), linalg::bicgstab_solver (krylov/bicgstab.cpp: */
/* 64-231) and linalg::gmres_solver (krylov/gmres.cpp:46-530).               */
/* ------------------------------------------------------------------------ */
/* x = rhs / diag elementwise: jacobi_precond::solve (host_math.cpp:53-64). */
int lal_jacobi_solve_f64(lal_handle h, int n, const double* rhs, const double* diag, double* x);

/* Preconditioner kinds of the Krylov solvers below. */
#define LAL_PRECOND_NONE 0
#define LAL_PRECOND_JACOBI 1 /* first diagonal entry of each row       */
#define LAL_PRECOND_SAAMG 2  /* one V-cycle of hier (opts->n1, n2 sweeps) */

/* BiCGSTAB on the square CSR matrix A (device arrays, descr from
 * lal_csrmv_analysis or NULL); x is the initial guess on entry.  Convergence
 * on the inf-norm residual (iter_control); info->iterations is the
 * reference's return value.  The breakdown branch (t.r or t.t exactly zero)
 * adds alpha p to x and stops, as bicgstab.cpp does.  hier is needed only for
 * LAL_PRECOND_SAAMG and must be built for A. */
int lal_bicgstab_solve(lal_handle h, lal_csrmv_descr descr, int m, int nnz, const int* row_ptr,
                       const int* col_ind, const double* val, const double* b, double* x,
                       int precond, lal_hierarchy hier, const lal_solve_options* opts,
                       lal_solve_info* info);
/* Restarted, left-preconditioned GMRES(restart) (restart clamped to
 * [1, m-1]); convergence on the 2-norm of the preconditioned residual;
 * info->iterations counts restart cycles as gmres.cpp does and
 * info->outer_iterations the Arnoldi steps (preconditioned products) taken.  The
 * preconditioned Arnoldi vector is kept and every work vector sized, where
 * gmres.cpp drops the one and leaves the others unsized. */
int lal_gmres_solve(lal_handle h, lal_csrmv_descr descr, int m, int nnz, const int* row_ptr,
                    const int* col_ind, const double* val, const double* b, double* x,
                    int restart, int precond, lal_hierarchy hier, const lal_solve_options* opts,
                    lal_solve_info* info);

/* ------------------------------------------------------------------------ */
/* Adjoint (no reference counterpart): for loss L(x) with x = A^-1 b and A   */
/* SPD, dL/db = lambda with A lambda = g, and dL/dA_ij = -lambda_i x_j over  */
/* A's pattern.                                                              */
/* ------------------------------------------------------------------------ */
/* A lambda = g with the hierarchy built for A: A is SPD, so the adjoint
 * system is the forward one and its preconditioner is reused.  lambda is
 * overwritten (zero initial guess).  mixed = 0: FP64 PCG; 1: FP32-inner /
 * FP64-outer refinement; 2: FP64 PCG with the FP32 V-cycle; 3: FP32 PCG with
 * FP64 reliable updates, as lal_pcg_solve_reliable (1-3 need a hierarchy
 * built with build_f32). */
int lal_adjoint_solve(lal_handle h, lal_hierarchy hier, const double* g, double* lambda,
                      int mixed, const lal_solve_options* opts, lal_solve_info* info);
/* vals[k] = scale * u[row(k)] * v[col(k)] over the CSR pattern; with
 * scale = -1, u = lambda, v = x this is dL/dA_ij = -lambda_i x_j. */
int lal_sddmm_f64(lal_handle h, int m, int nnz, const int* row_ptr, const int* col_ind,
                  double scale, const double* u, const double* v, double* vals);

#ifdef __cplusplus
}
#endif
#endif
