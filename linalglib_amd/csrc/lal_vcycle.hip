// V-cycle kernels: Jacobi smoothing (from zero, fused with the residual, and
// general sweeps), the coarse-level solves, and the PCG vector updates.
//
// Smoother arithmetic follows iterative_solvers/classic/jacobi.cpp:43-75:
// sigma = sigma + a_jk*xold_k over k != j in CSR order (two roundings per
// term), x_j = (b_j - sigma)/a_jj with a_jj the row's last diagonal entry.  A
// sweep from x = 0 gives sigma = +0 exactly, so it is x_j = b_j/a_jj; the
// fused kernel recomputes the neighbours' x_k = b_k/a_kk on the fly, which is
// bit-identical to smoothing first and forming r = b - A x afterwards.  These
// row-block kernels are the exact path; the fast path (lal_rowvec.hip) forms
// the same sweep as x + (b - A x) / a_ii.
//
// PCG reductions never use atomics or device-scope fences (on a multi-XCD
// part a device-scope release writes the L2 back, which costs more than the
// reduction itself).  Each reduction kernel writes one partial per block;
// the kernels that need the scalar re-reduce the partials themselves, every
// block in the same fixed order, so all blocks see bitwise the same value.
#include "lal_vcycle.h"

namespace
{
    using lal::SPMV_THREADS;
    constexpr int RB = lal::PCG_BLOCKS;

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

    // a_ii as the Jacobi sweep sees it (the last diagonal entry of the row)
    // and its reciprocal for the fast path.
    template <typename T>
    __global__ void __launch_bounds__(256)
        diag_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                    const T* __restrict__ val, T* __restrict__ d, T* __restrict__ dinv)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        T a = 0;
        for(int j = rp[i]; j < rp[i + 1]; j++)
            if(ci[j] == i)
                a = val[j];
        d[i]    = a;
        dinv[i] = (T)1 / a;
    }

    // diag_kernel with 64 lanes per row, for long rows: the row's last
    // diagonal entry is the group's largest matching position.
    template <typename T>
    __global__ void __launch_bounds__(256)
        diag_wave_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                         const T* __restrict__ val, T* __restrict__ d, T* __restrict__ dinv)
    {
        const int lane = (int)threadIdx.x & 63;
        const int i    = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
        if(i >= m)
            return;
        int last = -1;
        for(int j = rp[i] + lane; j < rp[i + 1]; j += 64)
            if(ci[j] == i)
                last = j;
        last = lal::wave_max(last);
        if(lane == 0)
        {
            const T a = last >= 0 ? val[last] : (T)0;
            d[i]      = a;
            dinv[i]   = (T)1 / a;
        }
    }

    // Coarsest level with one unknown: x = b / a (amg.cpp:211-216).
    template <typename T>
    __global__ void scalar_solve_kernel(const T* __restrict__ val, const T* __restrict__ b,
                                        T* __restrict__ x)
    {
        if(threadIdx.x == 0 && blockIdx.x == 0)
            x[0] = b[0] / val[0];
    }

    // 1000 in-place Gauss-Seidel sweeps from x = 0 (amg.cpp:236-264), one lane.
    template <typename T>
    __global__ void __launch_bounds__(64)
        coarse_gs_kernel(int n, const int* __restrict__ rp, const int* __restrict__ ci,
                         const T* __restrict__ val, const T* __restrict__ b, T* __restrict__ x,
                         int sweeps, int zero)
    {
        // The sweeps are one lane's sequential loop; when the operator, x and
        // b fit, the workgroup first stages them in LDS so that the loop's
        // dependent loads are LDS reads instead of L2 round trips.
        constexpr int          CGS_LDS = 40 * 1024;
        __shared__ __align__(16) unsigned char s_mem[CGS_LDS];
        const int              nnz  = rp[n];
        const size_t           need = sizeof(T) * ((size_t)nnz + 2 * (size_t)n)
                            + sizeof(int) * ((size_t)n + 1 + (size_t)nnz);
        const int tid = threadIdx.x;
        if(need <= CGS_LDS)
        {
            T*   s_v  = reinterpret_cast<T*>(s_mem);
            T*   s_x  = s_v + nnz;
            T*   s_b  = s_x + n;
            int* s_rp = reinterpret_cast<int*>(s_b + n);
            int* s_ci = s_rp + n + 1;
            for(int k = tid; k < nnz; k += blockDim.x)
            {
                s_v[k]  = val[k];
                s_ci[k] = ci[k];
            }
            for(int j = tid; j < n; j += blockDim.x)
            {
                s_x[j] = zero ? T(0) : x[j];
                s_b[j] = b[j];
            }
            for(int j = tid; j <= n; j += blockDim.x)
                s_rp[j] = rp[j];
            __syncthreads();
            if(tid == 0)
                for(int it = 0; it < sweeps; it++)
                    for(int j = 0; j < n; j++)
                    {
                        T sigma = 0, ajj = 0;
                        for(int k = s_rp[j]; k < s_rp[j + 1]; k++)
                        {
                            const int col = s_ci[k];
                            const T   v   = s_v[k];
                            if(col != j)
                                sigma = sigma + v * s_x[col];
                            else
                                ajj = v;
                        }
                        s_x[j] = (s_b[j] - sigma) / ajj;
                    }
            __syncthreads();
            for(int j = tid; j < n; j += blockDim.x)
                x[j] = s_x[j];
            return;
        }
        if(tid != 0)
            return;
        if(zero)
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
    // OP 0: sum; OP 1: maximum that keeps a NaN from either side, so that a
    // NaN anywhere in a vector reaches the solver's norm (and stops it)
    // whatever the order of the combination.
    template <int OP>
    __device__ __forceinline__ double combine(double a, double b)
    {
        return OP == 0 ? a + b : ((a != a || b < a) ? a : b);
    }

    // Block total of one value per lane (all lanes receive it).
    template <int OP>
    __device__ double block_total(double v)
    {
        __shared__ double sh[SPMV_THREADS / 64];
        v = OP == 0 ? lal::wave_sum(v) : lal::wave_max_nan(v);
        __syncthreads();
        if((threadIdx.x & 63) == 0)
            sh[threadIdx.x >> 6] = v;
        __syncthreads();
        double t = sh[0];
#pragma unroll
        for(int w = 1; w < SPMV_THREADS / 64; w++)
            t = combine<OP>(t, sh[w]);
        return t;
    }

    // Sum of the RB partials of a previous reduction, identical in every block.
    __device__ double sum_partials(const double* __restrict__ part)
    {
        double v = 0.0;
#pragma unroll
        for(int k = threadIdx.x; k < RB; k += SPMV_THREADS)
            v += part[k];
        return block_total<0>(v);
    }

    // part[block] = partial of sum x_i*y_i (OP 0) or max |x_i| (OP 1); x and y
    // may differ in precision (FP32 preconditioned vector, FP64 residual).
    template <typename TX, typename TY, int OP>
    __global__ void __launch_bounds__(SPMV_THREADS)
        partial_kernel(int n, const TX* __restrict__ x, const TY* __restrict__ y,
                       double* __restrict__ part)
    {
        double    acc    = 0.0;
        const int stride = gridDim.x * SPMV_THREADS;
        int       i      = blockIdx.x * SPMV_THREADS + threadIdx.x;
        for(; i + 3 * stride < n; i += 4 * stride)
        {
            TX a[4];
            TY b[4];
#pragma unroll
            for(int k = 0; k < 4; k++)
            {
                a[k] = x[i + k * stride];
                b[k] = OP == 0 ? y[i + k * stride] : (TY)a[k];
            }
#pragma unroll
            for(int k = 0; k < 4; k++)
                acc = OP == 0 ? acc + (double)a[k] * (double)b[k]
                              : combine<1>(acc, fabs((double)a[k]));
        }
        for(; i < n; i += stride)
            acc = OP == 0 ? acc + (double)x[i] * (double)y[i] : combine<1>(acc, fabs((double)x[i]));
        const double t = block_total<OP>(acc);
        if(threadIdx.x == 0)
            part[blockIdx.x] = t;
    }

    // One block: total of RB partials into a device slot and, when given, a
    // mapped host word (followed by a sequence number the host polls).
    template <int OP>
    __global__ void __launch_bounds__(SPMV_THREADS)
        finish_kernel(const double* __restrict__ part, double* __restrict__ slot,
                      double* host_val, unsigned* host_seq, unsigned seq)
    {
        double v = OP == 0 ? 0.0 : 0.0;
        for(int k = threadIdx.x; k < RB; k += SPMV_THREADS)
            v = combine<OP>(v, part[k]);
        const double t = block_total<OP>(v);
        if(threadIdx.x == 0)
        {
            *slot = t;
            if(host_val)
            {
                __hip_atomic_store(host_val, t, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_store(host_seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
            }
        }
    }

    // alpha = gamma / <q, p>; x += alpha p; r += (-1*alpha) q (cg.cpp:185-193);
    // part_r[block] = max |r_i|.
    // When x0 is given, also x0 = r * dinv: the first Jacobi iterate from zero
    // of the preconditioner applied next to r, in the preconditioner's
    // precision S; when rs is given, rs = (S) r, the preconditioner's input.
    template <typename T, typename S>
    __global__ void __launch_bounds__(SPMV_THREADS)
        update_xr_kernel(int n, const T* __restrict__ p, const T* __restrict__ q,
                         T* __restrict__ x, T* __restrict__ r,
                         const double* __restrict__ part_pq,
                         const double* __restrict__ part_gamma, double* __restrict__ part_r,
                         double* __restrict__ s_alpha, const S* __restrict__ dinv,
                         S* __restrict__ x0, S* __restrict__ rs, T rs_scale)
    {
        const int stride = gridDim.x * SPMV_THREADS;
        int       i      = blockIdx.x * SPMV_THREADS + threadIdx.x;
        const double pq    = sum_partials(part_pq);
        const double gamma = sum_partials(part_gamma);
        const double alpha = gamma / pq;
        if(blockIdx.x == 0 && threadIdx.x == 0)
            *s_alpha = alpha;
        const T a  = (T)alpha;
        const T ma = (T)(-1.0 * alpha);
        double  acc = 0.0;
        auto    upd = [&](T pi, T qi, T xi, T ri, S di, int at) {
            x[at]    = a == (T)1 ? pi + xi : a * pi + xi;
            T rn     = ma == (T)1 ? qi + ri : ma * qi + ri;
            r[at]    = rn;
            const S r_s = (S)(rn * rs_scale);
            if(rs)
                rs[at] = r_s;
            if(x0)
                x0[at] = r_s * di;
            acc = combine<1>(acc, fabs((double)rn));
        };
        for(; i + 3 * stride < n; i += 4 * stride)
        {
            T pp[4], qq[4], xx[4], rr[4];
            S dd[4];
#pragma unroll
            for(int k = 0; k < 4; k++)
            {
                pp[k] = p[i + k * stride];
                qq[k] = q[i + k * stride];
                xx[k] = x[i + k * stride];
                rr[k] = r[i + k * stride];
                dd[k] = x0 ? dinv[i + k * stride] : (S)0;
            }
#pragma unroll
            for(int k = 0; k < 4; k++)
                upd(pp[k], qq[k], xx[k], rr[k], dd[k], i + k * stride);
        }
        for(; i < n; i += stride)
            upd(p[i], q[i], x[i], r[i], x0 ? dinv[i] : (S)0, i);
        const double t = block_total<1>(acc);
        if(threadIdx.x == 0)
            part_r[blockIdx.x] = t;
    }

    // beta = gamma_new / gamma_old; p = 1*z + beta*p (cg.cpp:205-211); z may be
    // the FP32 output of an FP32 preconditioner.
    template <typename TZ, typename T>
    __global__ void __launch_bounds__(SPMV_THREADS)
        update_p_kernel(int n, const TZ* __restrict__ z, T* __restrict__ p,
                        const double* __restrict__ part_new,
                        const double* __restrict__ part_old)
    {
        const double g_new = sum_partials(part_new);
        const double g_old = sum_partials(part_old);
        const T      beta  = (T)(g_new / g_old);
        const int    stride = gridDim.x * SPMV_THREADS;
        int          i      = blockIdx.x * SPMV_THREADS + threadIdx.x;
        for(; i + 3 * stride < n; i += 4 * stride)
        {
            T zz[4], pp[4];
#pragma unroll
            for(int k = 0; k < 4; k++)
            {
                zz[k] = (T)z[i + k * stride];
                pp[k] = p[i + k * stride];
            }
#pragma unroll
            for(int k = 0; k < 4; k++)
                p[i + k * stride] = beta == (T)0 ? zz[k] : zz[k] + beta * pp[k];
        }
        for(; i < n; i += stride)
            p[i] = beta == (T)0 ? (T)z[i] : (T)z[i] + beta * p[i];
    }

    // x = b * dinv (first Jacobi iterate from zero)
    template <typename T>
    __global__ void __launch_bounds__(256)
        mul_kernel(int n, const T* __restrict__ b, const T* __restrict__ dinv,
                   T* __restrict__ x)
    {
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            x[i] = b[i] * dinv[i];
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
                     const int* ci, const T* v, const T* diag, const T* dinv, const T* xold,
                     const T* b, T* xnew)
    {
        if(m == 0)
            return LAL_OK;
        if(d && !d->exact)
        {
            rv_args a{m, d->nnz, d->by_nnz, d->d_blocks, d->nblocks, rp, ci, v, xold, dinv, b, xnew, nullptr, 0.0,
                      0.0, d->win};
            lal::use_c16(a, d);
            a.keep = d->keep;
            return rowvec_launch<T>(c, RV_JACOBI, a);
        }
        int nb = spmv_grid(d, m);
        jacobi_kernel<T, 0><<<nb, SPMV_THREADS, 0, c->stream>>>(
            d ? d->d_blocks : nullptr, nb, m, rp, ci, v, xold, nullptr, b, xnew, nullptr,
            1); // no fast descriptor: one lane per row
        LAL_LAUNCHED();
        (void)diag;
        return LAL_OK;
    }

    template <typename T>
    int jacobi0_residual(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                         const int* ci, const T* v, const T* diag, const T* dinv, const T* b,
                         T* x, T* r)
    {
        if(m == 0)
            return LAL_OK;
        if(d && !d->exact)
        {
            rv_args a{m, d->nnz, d->by_nnz, d->d_blocks, d->nblocks, rp, ci, v, nullptr, dinv, b, x, r, 0.0, 0.0, d->win};
            lal::use_c16(a, d);
            a.keep = d->keep;
            return rowvec_launch<T>(c, RV_JACOBI0, a);
        }
        int nb = spmv_grid(d, m);
        jacobi_kernel<T, 1><<<nb, SPMV_THREADS, 0, c->stream>>>(
            d ? d->d_blocks : nullptr, nb, m, rp, ci, v, nullptr, diag, b, x, r,
            1); // no fast descriptor: one lane per row
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int diag_inv(lal_context* c, int m, const int* rp, const int* ci, const T* v, T* d, T* dinv,
                 long nnz)
    {
        if(m == 0)
            return LAL_OK;
        // a lane per row walks the row alone: rows of 24 or more entries take
        // a wavefront each
        if(nnz >= 24L * m)
            diag_wave_kernel<T><<<(m + 3) / 4, 256, 0, c->stream>>>(m, rp, ci, v, d, dinv);
        else
            diag_kernel<T><<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, v, d, dinv);
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
        coarse_gs_kernel<T><<<1, 64, 0, c->stream>>>(n, rp, ci, v, b, x, 1000, 1);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int coarse_gs_sweeps(lal_context* c, int n, const int* rp, const int* ci, const double* v,
                         const double* b, double* x, int sweeps)
    {
        coarse_gs_kernel<double><<<1, 64, 0, c->stream>>>(n, rp, ci, v, b, x, sweeps, 0);
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

    template <typename TX, typename TY>
    int dot_partials(lal_context* c, int n, const TX* x, const TY* y, double* part)
    {
        partial_kernel<TX, TY, 0><<<RB, SPMV_THREADS, 0, c->stream>>>(n, x, y, part);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int norm_partials(lal_context* c, int n, const T* x, double* part)
    {
        partial_kernel<T, T, 1><<<RB, SPMV_THREADS, 0, c->stream>>>(n, x, x, part);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int finish_partials(lal_context* c, int op, const double* part, double* slot,
                        double* host_val, unsigned* host_seq, unsigned seq)
    {
        if(op == 0)
            finish_kernel<0><<<1, SPMV_THREADS, 0, c->stream>>>(part, slot, host_val, host_seq,
                                                                seq);
        else
            finish_kernel<1><<<1, SPMV_THREADS, 0, c->stream>>>(part, slot, host_val, host_seq,
                                                                seq);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T, typename S>
    int update_xr(lal_context* c, int n, const T* p, const T* q, T* x, T* r,
                  const double* part_pq, const double* part_gamma, double* part_r,
                  double* s_alpha, const S* dinv, S* x0, S* rs, double rs_scale)
    {
        update_xr_kernel<T, S><<<RB, SPMV_THREADS, 0, c->stream>>>(
            n, p, q, x, r, part_pq, part_gamma, part_r, s_alpha, dinv, x0, rs, (T)rs_scale);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int mul_elem(lal_context* c, int n, const T* b, const T* dinv, T* x)
    {
        if(n == 0)
            return LAL_OK;
        mul_kernel<T><<<grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, b, dinv, x);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int restrict_init(lal_context* c, const lal_csrmv_info* d, int m, const int* rp,
                      const int* ci, const T* v, const T* r, const T* dinv_c, T* b_c, T* x_c)
    {
        if(m == 0)
            return LAL_OK;
        rv_args a{m, d->nnz, d->by_nnz, d->d_blocks, d->nblocks, rp, ci, v, r, dinv_c, nullptr, b_c, x_c, 1.0, 0.0, d->win};
        lal::use_c16(a, d);
        a.keep = d->keep;
        return rowvec_launch<T>(c, RV_RESTRICT, a);
    }

    template <typename TZ, typename T>
    int update_p(lal_context* c, int n, const TZ* z, T* p, const double* part_new,
                 const double* part_old)
    {
        update_p_kernel<TZ, T><<<RB, SPMV_THREADS, 0, c->stream>>>(n, z, p, part_new,
                                                                   part_old);
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

#define LAL_INST(T)                                                                          \
    template int jacobi_sweep<T>(lal_context*, const lal_csrmv_info*, int, const int*,       \
                                 const int*, const T*, const T*, const T*, const T*,         \
                                 const T*, T*);                                              \
    template int jacobi0_residual<T>(lal_context*, const lal_csrmv_info*, int, const int*,   \
                                     const int*, const T*, const T*, const T*, const T*, T*, \
                                     T*);                                                    \
    template int diag_inv<T>(lal_context*, int, const int*, const int*, const T*, T*, T*,   \
                             long);                                                          \
    template int coarse_scalar<T>(lal_context*, const T*, const T*, T*);                     \
    template int coarse_gs<T>(lal_context*, int, const int*, const int*, const T*, const T*, \
                              T*);                                                           \
    template int coarse_dense<T>(lal_context*, int, const double*, const T*, T*);            \
    template int dot_partials<T, T>(lal_context*, int, const T*, const T*, double*);         \
    template int norm_partials<T>(lal_context*, int, const T*, double*);                     \
    template int update_xr<T, T>(lal_context*, int, const T*, const T*, T*, T*,              \
                                 const double*, const double*, double*, double*, const T*,   \
                                 T*, T*, double);                                                    \
    template int update_p<T, T>(lal_context*, int, const T*, T*, const double*,              \
                                const double*);                                              \
    template int mul_elem<T>(lal_context*, int, const T*, const T*, T*);                     \
    template int restrict_init<T>(lal_context*, const lal_csrmv_info*, int, const int*,      \
                                  const int*, const T*, const T*, const T*, T*, T*);
    LAL_INST(double)
    LAL_INST(float)
#undef LAL_INST
    // FP64 Krylov vectors with an FP32 preconditioner
    template int dot_partials<float, double>(lal_context*, int, const float*, const double*,
                                             double*);
    template int update_xr<double, float>(lal_context*, int, const double*, const double*,
                                          double*, double*, const double*, const double*,
                                          double*, double*, const float*, float*, float*,
                                          double);
    template int update_p<float, double>(lal_context*, int, const float*, double*,
                                         const double*, const double*);
}

extern "C" int lal_sddmm_f64(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                             double scale, const double* u, const double* v, double* vals)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && u && (nnz == 0 || (ci && v && vals))));
    return lal::sddmm(c, m, rp, ci, scale, u, v, vals);
}
