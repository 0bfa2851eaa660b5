// Internal launchers for the smoothers, coarse solves and PCG updates.
#pragma once

#include "lal_rowvec.h"

namespace lal
{
    // Device scalar slots of one solver (an array of LAL_SOLVER_SLOTS doubles).
    enum
    {
        S_GAMMA  = 0, // <z, r>
        S_PQ     = 1, // <A p, p>
        S_ALPHA  = 2,
        S_BETA   = 3,
        S_RNORM  = 4, // inf-norm of the recursively updated residual
        S_RNORM0 = 5,
        S_TMP    = 6,
        LAL_SOLVER_SLOTS = 8
    };

    int csrmv_f64(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz, double alpha,
                  const int* rp, const int* ci, const double* v, const double* x, double beta,
                  double* y);
    int csrmv_f32(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz, float alpha,
                  const int* rp, const int* ci, const float* v, const float* x, float beta,
                  float* y);
    int residual_f64(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     const int* rp, const int* ci, const double* v, const double* x,
                     const double* b, double* r);
    int residual_f32(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     const int* rp, const int* ci, const float* v, const float* x,
                     const float* b, float* r);
    // trusted: the pattern was built by this library (a hierarchy operator),
    // so the column-range check is skipped
    int csrmv_analysis_dev(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
                           int exact, lal_csrmv_info** out, bool trusted = false);

    inline int csrmv(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     double alpha, const int* rp, const int* ci, const double* v,
                     const double* x, double beta, double* y)
    {
        return csrmv_f64(c, d, m, n, nnz, alpha, rp, ci, v, x, beta, y);
    }
    inline int csrmv(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     double alpha, const int* rp, const int* ci, const float* v, const float* x,
                     double beta, float* y)
    {
        return csrmv_f32(c, d, m, n, nnz, (float)alpha, rp, ci, v, x, (float)beta, y);
    }

    // Grid of the PCG reduction and update kernels (one partial per block).
    constexpr int PCG_BLOCKS = 512;

    template <typename T>
    int jacobi_sweep(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                     const int* ci, const T* v, const T* diag, const T* dinv, const T* xold,
                     const T* b, T* xnew);
    template <typename T>
    int jacobi0_residual(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                         const int* ci, const T* v, const T* diag, const T* dinv, const T* b,
                         T* x, T* r);
    // d = the last diagonal entry of each row, dinv = 1/d.
    template <typename T>
    // nnz: the operator's nonzeros, which pick a lane or a wavefront per row
    int diag_inv(lal_context* c, int m, const int* rp, const int* ci, const T* v, T* d, T* dinv,
                 long nnz = 0);
    template <typename T>
    int coarse_scalar(lal_context* c, const T* v, const T* b, T* x);
    template <typename T>
    int coarse_gs(lal_context* c, int n, const int* rp, const int* ci, const T* v, const T* b,
                  T* x);
    template <typename T>
    int coarse_dense(lal_context* c, int n, const double* inv, const T* b, T* x);
    // `sweeps` sequential Gauss-Seidel sweeps on x as given (one lane).
    int coarse_gs_sweeps(lal_context* c, int n, const int* rp, const int* ci, const double* v,
                         const double* b, double* x, int sweeps);
    // PCG_BLOCKS partials of <x, y> / max|x| (FP64 accumulation).
    template <typename TX, typename TY>
    int dot_partials(lal_context* c, int n, const TX* x, const TY* y, double* part);
    template <typename T>
    int norm_partials(lal_context* c, int n, const T* x, double* part);
    // op 0: sum, 1: max of the partials into *slot (device) and, when host_val
    // is given (mapped pinned memory), into *host_val then *host_seq = seq.
    int finish_partials(lal_context* c, int op, const double* part, double* slot,
                        double* host_val, unsigned* host_seq, unsigned seq);
    // x += alpha p, r -= alpha q, partial max|r|; with x0 also
    // x0 = (S)(r * rs_scale) * dinv, with rs also rs = (S)(r * rs_scale) (the
    // preconditioner's precision S; rs_scale = 1 leaves r as it is).
    template <typename T, typename S>
    int update_xr(lal_context* c, int n, const T* p, const T* q, T* x, T* r,
                  const double* part_pq, const double* part_gamma, double* part_r,
                  double* s_alpha, const S* dinv, S* x0, S* rs, double rs_scale = 1.0);
    // x = b * dinv
    template <typename T>
    int mul_elem(lal_context* c, int n, const T* b, const T* dinv, T* x);
    // b_c = R r and x_c = b_c * dinv_c (fast path only: needs a descriptor)
    template <typename T>
    int restrict_init(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                      const int* ci, const T* v, const T* r, const T* dinv_c, T* b_c, T* x_c);
    template <typename TZ, typename T>
    int update_p(lal_context* c, int n, const TZ* z, T* p, const double* part_new,
                 const double* part_old);
    int scale_cast_f64_f32(lal_context* c, int n, const double* x, double scale, float* y);
    int axpy_f32_f64(lal_context* c, int n, double alpha, const float* d, double* x);
}
