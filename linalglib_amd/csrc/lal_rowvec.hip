// Staged-product CSR kernels (see lal_rowvec.h).
#include "lal_rowvec.h"

namespace
{
    using lal::SPMV_THREADS;

    template <typename T>
    __device__ __forceinline__ T stream_load(const T* p)
    {
        return __builtin_nontemporal_load(p);
    }

    template <typename T, int MODE>
    __device__ __forceinline__ T gather(const T* __restrict__ x, const T* __restrict__ d,
                                        const T* __restrict__ b, int c)
    {
        if(MODE == lal::RV_JACOBI0)
            return b[c] * d[c];
        return x[c];
    }

    template <typename T, int MODE>
    __global__ void __launch_bounds__(SPMV_THREADS)
        staged_kernel(int m, const int* __restrict__ blocks, int nb, const int* __restrict__ rp,
                      const int* __restrict__ ci, const T* __restrict__ val,
                      const T* __restrict__ x, const T* __restrict__ d,
                      const T* __restrict__ b, T alpha, T beta, T* __restrict__ y,
                      T* __restrict__ r)
    {
        constexpr int CAP = 2048;
        __shared__ T  s_p[CAP];

        const int tid  = threadIdx.x;
        const int tile = lal::xcd_tile(blockIdx.x, nb);
        int       r0, r1;
        lal::block_rows(blocks, m, tile, &r0, &r1);
        const int  nrows  = r1 - r0;
        const int  g      = lal::spmv_group(nrows, 0);
        const int  lrow   = tid / g;
        const int  lig    = tid & (g - 1);
        const bool active = lrow < nrows;
        const int  row    = r0 + lrow;
        const int  base   = rp[r0];
        const int  end    = rp[r1];
        const int  rs     = active ? rp[row] : 0;
        const int  re     = active ? rp[row + 1] : 0;

        T acc = 0;
        for(int cs = base; cs < end; cs += CAP)
        {
            const int  len = min(CAP, end - cs);
            const int* cc  = ci + cs;
            const T*   vv  = val + cs;
            int        k   = tid;
            for(; k + 3 * SPMV_THREADS < len; k += 4 * SPMV_THREADS)
            {
                int c[4];
                T   v[4], xx[4];
#pragma unroll
                for(int u = 0; u < 4; u++)
                {
                    c[u] = stream_load(cc + k + u * SPMV_THREADS);
                    v[u] = stream_load(vv + k + u * SPMV_THREADS);
                }
#pragma unroll
                for(int u = 0; u < 4; u++)
                    xx[u] = gather<T, MODE>(x, d, b, c[u]);
#pragma unroll
                for(int u = 0; u < 4; u++)
                    s_p[k + u * SPMV_THREADS] = v[u] * xx[u];
            }
            for(; k < len; k += SPMV_THREADS)
                s_p[k] = stream_load(vv + k) * gather<T, MODE>(x, d, b, stream_load(cc + k));
            __syncthreads();
            if(active)
            {
                const int a = max(rs, cs) - cs;
                const int e = min(re, cs + len) - cs;
                for(int j = a + lig; j < e; j += g)
                    acc += s_p[j];
            }
            __syncthreads();
        }
        if(g > 1)
            acc = lal::group_sum(acc, g);
        if(!active || lig != 0)
            return;
        if(MODE == lal::RV_SPMV)
            y[row] = beta == (T)0 ? alpha * acc : fma(alpha, acc, beta * y[row]);
        else if(MODE == lal::RV_RESTRICT)
        {
            y[row] = acc;
            r[row] = acc * d[row];
        }
        else if(MODE == lal::RV_RESIDUAL)
            y[row] = b[row] - acc;
        else if(MODE == lal::RV_JACOBI)
            y[row] = x[row] + (b[row] - acc) * d[row];
        else
        {
            const T bi = b[row];
            y[row]     = bi * d[row];
            r[row]     = bi - acc;
        }
    }

    template <typename T, int MODE>
    int launch(lal_context* c, const lal::rv_args& a)
    {
        const int nb = a.blocks ? a.nb : (a.m + SPMV_THREADS - 1) / SPMV_THREADS;
        if(nb == 0)
            return LAL_OK;
        staged_kernel<T, MODE><<<nb, SPMV_THREADS, 0, c->stream>>>(
            a.m, a.blocks, nb, a.rp, a.ci, (const T*)a.val, (const T*)a.x, (const T*)a.d,
            (const T*)a.b, (T)a.alpha, (T)a.beta, (T*)a.y, (T*)a.r);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

namespace lal
{
    template <typename T>
    int rowvec_launch(lal_context* c, int mode, const rv_args& a)
    {
        switch(mode)
        {
        case RV_SPMV: return launch<T, RV_SPMV>(c, a);
        case RV_RESTRICT: return launch<T, RV_RESTRICT>(c, a);
        case RV_RESIDUAL: return launch<T, RV_RESIDUAL>(c, a);
        case RV_JACOBI: return launch<T, RV_JACOBI>(c, a);
        default: return launch<T, RV_JACOBI0>(c, a);
        }
    }

    template int rowvec_launch<double>(lal_context*, int, const rv_args&);
    template int rowvec_launch<float>(lal_context*, int, const rv_args&);
}
