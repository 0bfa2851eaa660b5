// V-cycle kernels: Jacobi smoothing (from zero, fused with the residual, and
// general sweeps), the coarse-level solves, and the fused PCG vector updates
// with deterministic reductions.
//
// Smoother arithmetic follows iterative_solvers/classic/jacobi.cpp:43-75:
// sigma = sigma + a_jk*xold_k over k != j in CSR order (two roundings per
// term), x_j = (b_j - sigma)/a_jj with a_jj the row's last diagonal entry.  A
// sweep from x = 0 gives sigma = +0 exactly, so it is x_j = b_j/a_jj; the
// fused kernel recomputes the neighbours' x_k = b_k/a_kk on the fly, which is
// bit-identical to smoothing first and forming r = b - A x afterwards.
#include "lal_vcycle.h"

namespace
{
    using lal::SPMV_THREADS;

    // MODE 0: x_new = Jacobi sweep of xold ; MODE 1: x = b/d and r = b - A x
    template <typename T, int MODE>
    __global__ void __launch_bounds__(SPMV_THREADS)
        jacobi_kernel(const int* __restrict__ blocks, int nb, int m, const int* __restrict__ rp,
                      const int* __restrict__ ci, const T* __restrict__ val,
                      const T* __restrict__ xold, const T* __restrict__ d,
                      const T* __restrict__ b, T* __restrict__ xnew, T* __restrict__ r,
                      int exact)
    {
        constexpr int CAP = lal::spmv_cap<T>::value;
        __shared__ T   s_v[CAP];
        __shared__ T   s_x[CAP];
        __shared__ int s_c[MODE == 0 ? CAP : 1];
        __shared__ int s_rp[SPMV_THREADS + 1];

        const int tid  = threadIdx.x;
        const int tile = lal::xcd_tile(blockIdx.x, nb);
        int       r0, r1;
        lal::block_rows(blocks, m, tile, &r0, &r1);
        const int nrows = r1 - r0;
        for(int i = tid; i <= nrows; i += SPMV_THREADS)
            s_rp[i] = rp[r0 + i];
        __syncthreads();

        const int  base   = s_rp[0];
        const int  end    = s_rp[nrows];
        const int  g      = lal::spmv_group(nrows, exact);
        const int  lrow   = tid / g;
        const int  lig    = tid & (g - 1);
        const bool active = lrow < nrows;
        const int  row    = r0 + lrow;
        const int  rs     = active ? s_rp[lrow] : 0;
        const int  re     = active ? s_rp[lrow + 1] : 0;

        T sigma = 0, ajj = 0;
        for(int cs = base; cs < end; cs += CAP)
        {
            const int len = min(CAP, end - cs);
#pragma unroll 4
            for(int k = tid; k < len; k += SPMV_THREADS)
            {
                const int c = ci[cs + k];
                s_v[k]      = val[cs + k];
                if(MODE == 0)
                {
                    s_c[k] = c;
                    s_x[k] = xold[c];
                }
                else
                    s_x[k] = b[c] / d[c];
            }
            __syncthreads();
            if(active)
            {
                const int a = max(rs, cs) - cs;
                const int e = min(re, cs + len) - cs;
                for(int j = a + lig; j < e; j += g)
                {
                    if(MODE == 0)
                    {
                        if(s_c[j] != row)
                            sigma = sigma + s_v[j] * s_x[j];
                        else
                            ajj = s_v[j];
                    }
                    else
                        sigma = sigma + s_v[j] * s_x[j];
                }
            }
            __syncthreads();
        }
        if(g > 1)
        {
            sigma = lal::group_sum(sigma, g);
            if(MODE == 0)
                ajj = lal::group_sum(ajj, g);
        }
        if(active && lig == 0)
        {
            if(MODE == 0)
                xnew[row] = (b[row] - sigma) / ajj;
            else
            {
                xnew[row] = b[row] / d[row];
                r[row]    = b[row] - sigma;
            }
        }
    }

    // a_ii as the Jacobi sweep sees it: the last diagonal entry of the row.
    template <typename T>
    __global__ void __launch_bounds__(256)
        last_diag_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                         const T* __restrict__ val, T* __restrict__ d)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        T a = 0;
        for(int j = rp[i]; j < rp[i + 1]; j++)
            if(ci[j] == i)
                a = val[j];
        d[i] = a;
    }

    // Coarsest level with one unknown: x = b / a (amg.cpp:211-216).
    template <typename T>
    __global__ void scalar_solve_kernel(const T* __restrict__ val, const T* __restrict__ b,
                                        T* __restrict__ x)
    {
        if(threadIdx.x == 0 && blockIdx.x == 0)
            x[0] = b[0] / val[0];
    }

    // 1000 in-place Gauss-Seidel sweeps from x = 0 (amg.cpp:236-264), one
    // lane, the matrix staged in LDS when it fits.
    template <typename T>
    __global__ void __launch_bounds__(64)
        coarse_gs_kernel(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                         const T* __restrict__ val, const T* __restrict__ b, T* __restrict__ x,
                         int sweeps)
    {
        if(threadIdx.x != 0)
            return;
        for(int j = 0; j < n; j++)
            x[j] = 0;
        for(int it = 0; it < sweeps; it++)
            for(int j = 0; j < n; j++)
            {
                T sigma = 0, ajj = 0;
                for(int k = rp[j]; k < rp[j + 1]; k++)
                {
                    int col = ci[k];
                    T   v   = val[k];
                    if(col != j)
                        sigma = sigma + v * x[col];
                    else
                        ajj = v;
                }
                x[j] = (b[j] - sigma) / ajj;
            }
    }

    // Dense coarse solve x = Ainv b with the inverse kept in FP64.
    template <typename T>
    __global__ void __launch_bounds__(256)
        dense_apply_kernel(int n, const double* __restrict__ inv, const T* __restrict__ b,
                           T* __restrict__ x)
    {
        const int row  = blockIdx.x * 4 + (threadIdx.x >> 6);
        const int lane = threadIdx.x & 63;
        if(row >= n)
            return;
        double s = 0.0;
        for(int k = lane; k < n; k += 64)
            s = fma(inv[(size_t)row * n + k], (double)b[k], s);
        s = lal::wave_sum(s);
        if(lane == 0)
            x[row] = (T)s;
    }

    // ---------------------------------------------------- PCG vector updates
    // Reduction epilogues run by the last block on the device scalars.
    enum
    {
        EPI_STORE = 0, // s[out] = v
        EPI_ALPHA = 1, // s[PQ] = v ; s[ALPHA] = s[GAMMA] / v
        EPI_BETA  = 2  // s[BETA] = v / s[GAMMA] ; s[GAMMA] = v
    };

    template <int OP>
    __device__ __forceinline__ double red_combine(double a, double b)
    {
        return OP == 0 ? a + b : ((b < a) ? a : b);
    }

    // Block-level finish of a deterministic two-stage reduction: returns true
    // in thread 0 of the last block, with the total in *total.
    template <int OP>
    __device__ bool reduce_finish(double acc, double* __restrict__ part,
                                  unsigned* __restrict__ ticket, double* total)
    {
        __shared__ double sh[SPMV_THREADS / 64];
        __shared__ bool   last;
        const int         tid = threadIdx.x;
        acc                   = OP == 0 ? lal::wave_sum(acc) : lal::wave_max(acc);
        if((tid & 63) == 0)
            sh[tid >> 6] = acc;
        __syncthreads();
        if(tid == 0)
        {
            double t = sh[0];
            for(int w = 1; w < SPMV_THREADS / 64; w++)
                t = red_combine<OP>(t, sh[w]);
            part[blockIdx.x] = t;
            __threadfence();
            unsigned k = atomicAdd(ticket, 1u);
            last       = (k == gridDim.x - 1);
        }
        __syncthreads();
        if(!last)
            return false;
        __threadfence();
        double v = 0.0;
        for(int k = tid; k < (int)gridDim.x; k += SPMV_THREADS)
            v = red_combine<OP>(
                v, __hip_atomic_load(&part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        v = OP == 0 ? lal::wave_sum(v) : lal::wave_max(v);
        __syncthreads();
        if((tid & 63) == 0)
            sh[tid >> 6] = v;
        __syncthreads();
        if(tid != 0)
            return false;
        double t = sh[0];
        for(int w = 1; w < SPMV_THREADS / 64; w++)
            t = red_combine<OP>(t, sh[w]);
        *total  = t;
        *ticket = 0u;
        return true;
    }

    // dot(x, y) accumulated in FP64 with an epilogue on the solver scalars.
    template <typename T, int EPI>
    __global__ void __launch_bounds__(SPMV_THREADS)
        dot_epi_kernel(int n, const T* __restrict__ x, const T* __restrict__ y,
                       double* __restrict__ part, unsigned* __restrict__ ticket,
                       double* __restrict__ s, int out)
    {
        double acc = 0.0;
        for(int i = blockIdx.x * SPMV_THREADS + threadIdx.x; i < n;
            i += gridDim.x * SPMV_THREADS)
            acc = acc + (double)x[i] * (double)y[i];
        double v;
        if(!reduce_finish<0>(acc, part, ticket, &v))
            return;
        if(EPI == EPI_STORE)
            s[out] = v;
        else if(EPI == EPI_ALPHA)
        {
            s[lal::S_PQ]    = v;
            s[lal::S_ALPHA] = s[lal::S_GAMMA] / v;
        }
        else
        {
            s[lal::S_BETA]  = v / s[lal::S_GAMMA];
            s[lal::S_GAMMA] = v;
        }
    }

    // x += alpha p ; r += (-1*alpha) q ; s[RNORM] = max |r_i|   (cg.cpp:189-195)
    template <typename T>
    __global__ void __launch_bounds__(SPMV_THREADS)
        update_xr_kernel(int n, const T* __restrict__ p, const T* __restrict__ q,
                         T* __restrict__ x, T* __restrict__ r, double* __restrict__ part,
                         unsigned* __restrict__ ticket, double* __restrict__ s)
    {
        const T alpha  = (T)s[lal::S_ALPHA];
        const T malpha = (T)(-1.0 * s[lal::S_ALPHA]);
        double  acc    = 0.0;
        for(int i = blockIdx.x * SPMV_THREADS + threadIdx.x; i < n;
            i += gridDim.x * SPMV_THREADS)
        {
            x[i]     = alpha == (T)1 ? p[i] + x[i] : alpha * p[i] + x[i];
            T ri     = malpha == (T)1 ? q[i] + r[i] : malpha * q[i] + r[i];
            r[i]     = ri;
            double a = fabs((double)ri);
            acc      = (a < acc) ? acc : a;
        }
        double v;
        if(reduce_finish<1>(acc, part, ticket, &v))
            s[lal::S_RNORM] = v;
    }

    // p = 1*z + beta*p  (axpby, cg.cpp:210-211)
    template <typename T>
    __global__ void __launch_bounds__(256)
        update_p_kernel(int n, const T* __restrict__ z, T* __restrict__ p,
                        const double* __restrict__ s)
    {
        const T beta = (T)s[lal::S_BETA];
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            p[i] = beta == (T)0 ? z[i] : z[i] + beta * p[i];
    }

    // s[out] = max |x_i| (FP64 result for either precision)
    template <typename T>
    __global__ void __launch_bounds__(SPMV_THREADS)
        norm_inf_kernel(int n, const T* __restrict__ x, double* __restrict__ part,
                        unsigned* __restrict__ ticket, double* __restrict__ s, int out)
    {
        double acc = 0.0;
        for(int i = blockIdx.x * SPMV_THREADS + threadIdx.x; i < n;
            i += gridDim.x * SPMV_THREADS)
        {
            double a = fabs((double)x[i]);
            acc      = (a < acc) ? acc : a;
        }
        double v;
        if(reduce_finish<1>(acc, part, ticket, &v))
            s[out] = v;
    }

    // y32 = (float)(scale * x64): the refinement residual, normalised
    __global__ void __launch_bounds__(256)
        scale_cast_kernel(int n, const double* __restrict__ x, double scale,
                          float* __restrict__ y)
    {
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            y[i] = (float)(scale * x[i]);
    }

    // x64 += alpha * (double) d32
    __global__ void __launch_bounds__(256)
        axpy_f32_f64_kernel(int n, double alpha, const float* __restrict__ d,
                            double* __restrict__ x)
    {
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            x[i] = alpha * (double)d[i] + x[i];
    }

    __global__ void __launch_bounds__(256)
        sddmm_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci, double scale,
                     const double* __restrict__ u, const double* __restrict__ v,
                     double* __restrict__ out)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        const double su = scale * u[i];
        for(int j = rp[i]; j < rp[i + 1]; j++)
            out[j] = su * v[ci[j]];
    }
}

namespace lal
{
    template <typename T>
    int jacobi_sweep(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                     const int* ci, const T* v, const T* xold, const T* b, T* xnew)
    {
        if(m == 0)
            return LAL_OK;
        int nb = spmv_grid(d, m);
        jacobi_kernel<T, 0><<<nb, SPMV_THREADS, 0, c->stream>>>(
            d ? d->d_blocks : nullptr, nb, m, rp, ci, v, xold, nullptr, b, xnew, nullptr,
            d ? d->exact : 0);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int jacobi0_residual(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                         const int* ci, const T* v, const T* diag, const T* b, T* x, T* r)
    {
        if(m == 0)
            return LAL_OK;
        int nb = spmv_grid(d, m);
        jacobi_kernel<T, 1><<<nb, SPMV_THREADS, 0, c->stream>>>(
            d ? d->d_blocks : nullptr, nb, m, rp, ci, v, nullptr, diag, b, x, r,
            d ? d->exact : 0);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int last_diag(lal_context* c, int m, const int* rp, const int* ci, const T* v, T* d)
    {
        if(m == 0)
            return LAL_OK;
        last_diag_kernel<T><<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, v, d);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int coarse_scalar(lal_context* c, const T* v, const T* b, T* x)
    {
        scalar_solve_kernel<T><<<1, 64, 0, c->stream>>>(v, b, x);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int coarse_gs(lal_context* c, int n, const int* rp, const int* ci, const T* v, const T* b,
                  T* x)
    {
        coarse_gs_kernel<T><<<1, 64, 0, c->stream>>>(n, rp, ci, v, b, x, 1000);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int coarse_dense(lal_context* c, int n, const double* inv, const T* b, T* x)
    {
        dense_apply_kernel<T><<<(n + 3) / 4, 256, 0, c->stream>>>(n, inv, b, x);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    static int red_grid(int n)
    {
        return n <= 0 ? 1 : grid_for(n, SPMV_THREADS, LAL_RED_BLOCKS);
    }

    template <typename T>
    int dot_store(lal_context* c, int n, const T* x, const T* y, double* s, int out)
    {
        dot_epi_kernel<T, EPI_STORE><<<red_grid(n), SPMV_THREADS, 0, c->stream>>>(
            n, x, y, c->d_part, c->d_ticket, s, out);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int dot_alpha(lal_context* c, int n, const T* q, const T* p, double* s)
    {
        dot_epi_kernel<T, EPI_ALPHA><<<red_grid(n), SPMV_THREADS, 0, c->stream>>>(
            n, q, p, c->d_part, c->d_ticket, s, 0);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int dot_beta(lal_context* c, int n, const T* z, const T* r, double* s)
    {
        dot_epi_kernel<T, EPI_BETA><<<red_grid(n), SPMV_THREADS, 0, c->stream>>>(
            n, z, r, c->d_part, c->d_ticket, s, 0);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int update_xr(lal_context* c, int n, const T* p, const T* q, T* x, T* r, double* s)
    {
        update_xr_kernel<T><<<red_grid(n), SPMV_THREADS, 0, c->stream>>>(
            n, p, q, x, r, c->d_part, c->d_ticket, s);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int update_p(lal_context* c, int n, const T* z, T* p, const double* s)
    {
        update_p_kernel<T><<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, z, p, s);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int norm_inf_slot(lal_context* c, int n, const T* x, double* s, int out)
    {
        norm_inf_kernel<T><<<red_grid(n), SPMV_THREADS, 0, c->stream>>>(n, x, c->d_part,
                                                                       c->d_ticket, s, out);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int scale_cast_f64_f32(lal_context* c, int n, const double* x, double scale, float* y)
    {
        scale_cast_kernel<<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, x, scale, y);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int axpy_f32_f64(lal_context* c, int n, double alpha, const float* d, double* x)
    {
        axpy_f32_f64_kernel<<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, alpha, d, x);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int sddmm(lal_context* c, int m, const int* rp, const int* ci, double scale, const double* u,
              const double* v, double* out)
    {
        if(m == 0)
            return LAL_OK;
        sddmm_kernel<<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, scale, u, v, out);
        LAL_LAUNCHED();
        return LAL_OK;
    }

#define LAL_INST(T)                                                                         \
    template int jacobi_sweep<T>(lal_context*, const lal_csrmv_info*, int, const int*,      \
                                 const int*, const T*, const T*, const T*, T*);             \
    template int jacobi0_residual<T>(lal_context*, const lal_csrmv_info*, int, const int*,  \
                                     const int*, const T*, const T*, const T*, T*, T*);     \
    template int last_diag<T>(lal_context*, int, const int*, const int*, const T*, T*);     \
    template int coarse_scalar<T>(lal_context*, const T*, const T*, T*);                    \
    template int coarse_gs<T>(lal_context*, int, const int*, const int*, const T*, const T*, \
                              T*);                                                          \
    template int coarse_dense<T>(lal_context*, int, const double*, const T*, T*);           \
    template int dot_store<T>(lal_context*, int, const T*, const T*, double*, int);         \
    template int dot_alpha<T>(lal_context*, int, const T*, const T*, double*);              \
    template int dot_beta<T>(lal_context*, int, const T*, const T*, double*);               \
    template int update_xr<T>(lal_context*, int, const T*, const T*, T*, T*, double*);      \
    template int update_p<T>(lal_context*, int, const T*, T*, const double*);              \
    template int norm_inf_slot<T>(lal_context*, int, const T*, double*, int);
    LAL_INST(double)
    LAL_INST(float)
#undef LAL_INST
}

extern "C" int lal_sddmm_f64(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                             double scale, const double* u, const double* v, double* vals)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && u && (nnz == 0 || (ci && v && vals))));
    return lal::sddmm(c, m, rp, ci, scale, u, v, vals);
}
