// Internal launchers for the smoothers, coarse solves and PCG updates.
#pragma once

#include "lal_spmv.h"

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
    int csrmv_analysis_dev(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
                           int exact, lal_csrmv_info** out);

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

    template <typename T>
    int jacobi_sweep(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                     const int* ci, const T* v, const T* xold, const T* b, T* xnew);
    template <typename T>
    int jacobi0_residual(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                         const int* ci, const T* v, const T* diag, const T* b, T* x, T* r);
    template <typename T>
    int last_diag(lal_context* c, int m, const int* rp, const int* ci, const T* v, T* d);
    template <typename T>
    int coarse_scalar(lal_context* c, const T* v, const T* b, T* x);
    template <typename T>
    int coarse_gs(lal_context* c, int n, const int* rp, const int* ci, const T* v, const T* b,
                  T* x);
    template <typename T>
    int coarse_dense(lal_context* c, int n, const double* inv, const T* b, T* x);
    template <typename T>
    int dot_store(lal_context* c, int n, const T* x, const T* y, double* s, int out);
    template <typename T>
    int dot_alpha(lal_context* c, int n, const T* q, const T* p, double* s);
    template <typename T>
    int dot_beta(lal_context* c, int n, const T* z, const T* r, double* s);
    template <typename T>
    int update_xr(lal_context* c, int n, const T* p, const T* q, T* x, T* r, double* s);
    template <typename T>
    int update_p(lal_context* c, int n, const T* z, T* p, const double* s);
    template <typename T>
    int norm_inf_slot(lal_context* c, int n, const T* x, double* s, int out);
    int scale_cast_f64_f32(lal_context* c, int n, const double* x, double scale, float* y);
    int axpy_f32_f64(lal_context* c, int n, double alpha, const float* d, double* x);
}
