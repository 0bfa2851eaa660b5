// CSR SpMV, residual and diagonal extraction for gfx950.
//
// y = alpha*A*x + beta*y follows backend/host/host_matrix_vector.cpp:59-84
// (row sum s = fma(val, x[col], s), then beta == 0 ? alpha*s :
// fma(alpha, s, beta*y)); r = b - A*x follows :35-57 (s += val*x, two
// roundings).  The row-block layout and LDS staging are described in
// lal_spmv.h.
#include "lal_rowvec.h"

#include <algorithm>
#include <climits>

#include <vector>

namespace
{
    using lal::SPMV_THREADS;

    // MODE 0: y = alpha*A*x + beta*y ;  MODE 1: y = b - A*x
    template <typename T, typename ACC, int MODE>
    __global__ void __launch_bounds__(SPMV_THREADS)
        csrmv_kernel(const int* __restrict__ blocks, int nb, int m, const int* __restrict__ rp,
                     const int* __restrict__ ci, const T* __restrict__ val,
                     const T* __restrict__ x, ACC alpha, ACC beta, const T* __restrict__ b,
                     T* __restrict__ y, int exact)
    {
        constexpr int CAP = lal::spmv_cap<T>::value;
        __shared__ T  s_v[CAP];
        __shared__ T  s_x[CAP];
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
        const int  rs     = active ? s_rp[lrow] : 0;
        const int  re     = active ? s_rp[lrow + 1] : 0;

        ACC acc = 0;
        for(int cs = base; cs < end; cs += CAP)
        {
            const int len = min(CAP, end - cs);
#pragma unroll 4
            for(int k = tid; k < len; k += SPMV_THREADS)
            {
                int c  = ci[cs + k];
                s_v[k] = val[cs + k];
                s_x[k] = x[c];
            }
            __syncthreads();
            if(active)
            {
                const int a = max(rs, cs) - cs;
                const int e = min(re, cs + len) - cs;
                for(int j = a + lig; j < e; j += g)
                {
                    if(MODE == 0)
                        acc = fma((ACC)s_v[j], (ACC)s_x[j], acc);
                    else
                        acc = acc + (ACC)s_v[j] * (ACC)s_x[j];
                }
            }
            __syncthreads();
        }
        if(g > 1)
            acc = lal::group_sum(acc, g);
        if(active && lig == 0)
        {
            const int row = r0 + lrow;
            if(MODE == 0)
                y[row] = (T)(beta == (ACC)0 ? alpha * acc : fma(alpha, acc, beta * (ACC)y[row]));
            else
                y[row] = (T)((ACC)b[row] - acc);
        }
    }

    // Flags 1 in *bad if a column index lies outside [0, n): an invalid matrix
    // is refused before any gather can run off the end of x.
    __global__ void __launch_bounds__(256)
        check_cols_kernel(int nnz, int n, const int* __restrict__ ci, int* __restrict__ bad)
    {
        int k = blockIdx.x * 256 + threadIdx.x;
        int stride = gridDim.x * 256;
        int b = 0;
        for(; k < nnz; k += stride)
            b |= (unsigned)ci[k] >= (unsigned)n;
        if(__any(b) && (threadIdx.x & 63) == 0)
            atomicOr(bad, 1);
    }

    __global__ void __launch_bounds__(256)
        diagonal_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                        const double* __restrict__ val, double* __restrict__ d)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        for(int j = rp[i]; j < rp[i + 1]; j++)
            if(ci[j] == i)
            {
                d[i] = val[j];
                break;
            }
    }

    // One workgroup per row block: the block's column range, its base (the
    // smallest column) and every entry's offset from it as 16 bits; *fail
    // is set when a block spans more than 65535 columns.
    __global__ void __launch_bounds__(256)
        c16_kernel(const int* __restrict__ blocks, int nb, const int* __restrict__ ci,
                   int* __restrict__ cbase, unsigned short* __restrict__ c16, int* fail)
    {
        typedef int int4v __attribute__((ext_vector_type(4)));
        __shared__ int sh_mn[4], sh_mx[4];
        const int4v mt = reinterpret_cast<const int4v*>(blocks + lal::block_meta_offset(nb))[blockIdx.x];
        const int   e0 = mt.z, e1 = mt.w;
        int         mn = INT_MAX, mx = INT_MIN;
        for(int k = e0 + (int)threadIdx.x; k < e1; k += 256)
        {
            const int col = ci[k];
            mn            = min(mn, col);
            mx            = max(mx, col);
        }
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
        {
            mn = min(mn, __shfl_xor(mn, o, 64));
            mx = max(mx, __shfl_xor(mx, o, 64));
        }
        if((threadIdx.x & 63) == 0)
        {
            sh_mn[threadIdx.x >> 6] = mn;
            sh_mx[threadIdx.x >> 6] = mx;
        }
        __syncthreads();
        mn = min(min(sh_mn[0], sh_mn[1]), min(sh_mn[2], sh_mn[3]));
        mx = max(max(sh_mx[0], sh_mx[1]), max(sh_mx[2], sh_mx[3]));
        const int base = e1 > e0 ? mn : 0;
        if(threadIdx.x == 0)
        {
            cbase[blockIdx.x] = base;
            if(e1 > e0 && (long long)mx - mn > 65535)
                atomicOr(fail, 1);
        }
        for(int k = e0 + (int)threadIdx.x; k < e1; k += 256)
            c16[k] = (unsigned short)(ci[k] - base);
    }

    template <typename T, typename ACC, int MODE>
    int csrmv_launch(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz, ACC alpha,
                     const int* rp, const int* ci, const T* v, const T* x, ACC beta,
                     const T* b, T* y)
    {
        LAL_REQUIRE(c != nullptr && m >= 0 && n >= 0 && nnz >= 0);
        if(m == 0)
            return LAL_OK;
        LAL_REQUIRE(rp != nullptr && y != nullptr);
        LAL_REQUIRE(nnz == 0 || (ci != nullptr && v != nullptr && x != nullptr));
        if(d)
            LAL_REQUIRE(d->m == m && d->nnz == nnz);
        if(d && !d->exact)
        {
            lal::rv_args a{m,    d->nnz, d->by_nnz, d->d_blocks, d->nblocks,     rp,           ci, v, x, nullptr,
                           b,    y,           nullptr,        (double)alpha, (double)beta, d->win};
            a.keep = d->keep;
            lal::use_c16(a, d);
            return lal::rowvec_launch<T>(c, MODE == 0 ? lal::RV_SPMV : lal::RV_RESIDUAL, a);
        }
        // an exact descriptor, or none: one lane per row in CSR order (the
        // host loop's rounding)
        int nb = lal::spmv_grid(d, m);
        csrmv_kernel<T, ACC, MODE><<<nb, SPMV_THREADS, 0, c->stream>>>(
            d ? d->d_blocks : nullptr, nb, m, rp, ci, v, x, alpha, beta, b, y, 1);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

namespace lal
{
    // Greedy row packing: a block closes at 256 rows or when the next row
    // would push it past `cap` nonzeros; a row longer than cap is a block of
    // its own.  cap is 2048 (the kernels' LDS chunk) for large matrices; for
    // small ones it shrinks, down to 256, so that a launch still has about
    // 1024 workgroups: their time is one latency chain per workgroup, and
    // more, shorter chains spread over all CUs finish sooner.
    int csrmv_analysis_host(const int* h_rp, int m, std::vector<int>& blocks)
    {
        const long nnz = m > 0 ? h_rp[m] : 0;
        int        cap = spmv_cap<double>::value;
        while(cap > 256 && nnz / 1024 < cap)
            cap >>= 1;
        blocks.clear();
        blocks.push_back(0);
        int r = 0;
        while(r < m)
        {
            // the block ends before the first row r (at most SPMV_THREADS rows
            // on) with row_ptr[r + 1] - row_ptr[start] > cap; row_ptr is
            // non-decreasing, so that row is a binary search away
            const int       start = r;
            const int       lim   = std::min(m, start + SPMV_THREADS);
            const long long top   = (long long)h_rp[start] + cap;
            r = (int)(std::upper_bound(h_rp + start + 1, h_rp + lim + 1, top) - h_rp) - 1;
            if(r == start)
                r++;
            blocks.push_back(r);
        }
        return LAL_OK;
    }

    int check_columns(lal_context* c, int nnz, int n, const int* ci)
    {
        if(nnz == 0)
            return LAL_OK;
        if(ci == nullptr)
            return LAL_ERR_INVALID;
        LAL_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        check_cols_kernel<<<grid_for(nnz, 256, 4096), 256, 0, c->stream>>>(nnz, n, ci,
                                                                            c->d_flag);
        LAL_LAUNCHED();
        LAL_HIP(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        if(c->h_flag[0])
        {
            set_last_error("column index outside [0, n)");
            return LAL_ERR_INVALID;
        }
        return LAL_OK;
    }

    int csrmv_analysis_dev(lal_context* c, int m, int n, int nnz, const int* rp,
                           const int* ci, int exact, lal_csrmv_info** out, bool trusted)
    {
        if(!trusted)
            LAL_TRY(check_columns(c, nnz, n, ci));
        std::vector<int> h(m + 1);
        LAL_HIP(hipMemcpyAsync(h.data(), rp, sizeof(int) * (m + 1), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        if(h[0] != 0 || h[m] != nnz)
        {
            set_last_error("row_ptr does not start at 0 or end at nnz");
            return LAL_ERR_INVALID;
        }
        // (a pattern this library built is monotone by construction)
        for(int i = 0; i < m && !trusted; i++)
            if(h[i + 1] < h[i])
            {
                set_last_error("row_ptr is not monotone");
                return LAL_ERR_INVALID;
            }
        std::vector<int> blocks;
        csrmv_analysis_host(h.data(), m, blocks);
        auto* d    = new lal_csrmv_info;
        d->m       = m;
        d->nnz     = nnz;
        d->exact   = exact;
        d->by_nnz  = m != n;
        d->nblocks = (int)blocks.size() - 1;
        // the row offsets, then (16-byte aligned) one {r0, r1, row_ptr[r0],
        // row_ptr[r1]} quadruple per block: a block learns its rows and its
        // nonzero window from a single load
        const size_t meta_at = lal::block_meta_offset(d->nblocks);
        std::vector<int> packed(meta_at + 4 * (size_t)d->nblocks);
        std::copy(blocks.begin(), blocks.end(), packed.begin());
        int win = 4;
        for(int k = 0; k < d->nblocks; k++)
        {
            const int span = ((h[blocks[k + 1]] + 3) & ~3) - (h[blocks[k]] & ~3);
            win            = std::max(win, std::min(span, 2052));
            packed[meta_at + 4 * k + 0] = blocks[k];
            packed[meta_at + 4 * k + 1] = blocks[k + 1];
            packed[meta_at + 4 * k + 2] = h[blocks[k]];
            packed[meta_at + 4 * k + 3] = h[blocks[k + 1]];
        }
        d->win = win;
        if(hipMalloc(&d->d_blocks, sizeof(int) * packed.size()) != hipSuccess)
        {
            delete d;
            return LAL_ERR_OOM;
        }
        LAL_HIP(hipMemcpyAsync(d->d_blocks, packed.data(), sizeof(int) * packed.size(),
                               hipMemcpyHostToDevice, c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        *out = d;
        return LAL_OK;
    }

    int compact_columns(lal_context* c, lal_csrmv_info* d, int nnz, const int* ci)
    {
        if(!d || d->exact || nnz <= 0 || d->nblocks <= 0 || d->d_c16)
            return LAL_OK;
        unsigned short* c16   = nullptr;
        int*            cbase = nullptr;
        int*            fail  = c->d_flag; // the context's flag word, read back below
        auto release = [&]() {
            (void)hipFree(c16);
            (void)hipFree(cbase);
        };
        // rounded up to whole 16-byte lines: the kernels load four entries at a time
        if(hipMalloc(&c16, ((size_t)nnz * sizeof(unsigned short) + 15) & ~(size_t)15) !=
               hipSuccess ||
           hipMalloc(&cbase, sizeof(int) * d->nblocks) != hipSuccess)
        {
            release();
            return LAL_ERR_OOM;
        }
        int rc = hipMemsetAsync(fail, 0, sizeof(int), c->stream) == hipSuccess ? LAL_OK : LAL_ERR_HIP;
        if(rc == LAL_OK)
        {
            c16_kernel<<<d->nblocks, 256, 0, c->stream>>>(d->d_blocks, d->nblocks, ci, cbase, c16,
                                                         fail);
            if(hipGetLastError() != hipSuccess ||
               hipMemcpyAsync(c->h_flag, fail, sizeof(int), hipMemcpyDeviceToHost, c->stream) !=
                   hipSuccess ||
               hipStreamSynchronize(c->stream) != hipSuccess)
                rc = LAL_ERR_HIP;
        }
        const int h_fail = rc == LAL_OK ? c->h_flag[0] : 1;
        if(rc != LAL_OK || h_fail)
        {
            release();
            return rc;
        }
        d->d_c16   = c16;
        d->d_cbase = cbase;
        return LAL_OK;
    }

    int csrmv_f64(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz, double alpha,
                  const int* rp, const int* ci, const double* v, const double* x, double beta,
                  double* y)
    {
        return csrmv_launch<double, double, 0>(c, d, m, n, nnz, alpha, rp, ci, v, x, beta,
                                               nullptr, y);
    }

    int csrmv_f32(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz, float alpha,
                  const int* rp, const int* ci, const float* v, const float* x, float beta,
                  float* y)
    {
        return csrmv_launch<float, float, 0>(c, d, m, n, nnz, alpha, rp, ci, v, x, beta,
                                             nullptr, y);
    }

    int residual_f64(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     const int* rp, const int* ci, const double* v, const double* x,
                     const double* b, double* r)
    {
        LAL_REQUIRE(m == 0 || b != nullptr);
        return csrmv_launch<double, double, 1>(c, d, m, n, nnz, 0.0, rp, ci, v, x, 0.0, b, r);
    }

    int residual_f32(lal_context* c, const lal_csrmv_info* d, int m, int n, int nnz,
                     const int* rp, const int* ci, const float* v, const float* x,
                     const float* b, float* r)
    {
        LAL_REQUIRE(m == 0 || b != nullptr);
        return csrmv_launch<float, float, 1>(c, d, m, n, nnz, 0.0f, rp, ci, v, x, 0.0f, b, r);
    }

    int diagonal_f64(lal_context* c, int m, const int* rp, const int* ci, const double* v,
                     double* d)
    {
        if(m == 0)
            return LAL_OK;
        diagonal_kernel<<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, v, d);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

extern "C" int lal_csrmv_analysis(lal_handle c, int m, int n, int nnz, const int* row_ptr,
                                  const int* col_ind, int exact, lal_csrmv_descr* descr)
{
    LAL_REQUIRE(c != nullptr && descr != nullptr && m >= 0 && n >= 0 && nnz >= 0);
    LAL_REQUIRE(row_ptr != nullptr && (nnz == 0 || col_ind != nullptr));
    *descr = nullptr;
    return lal::csrmv_analysis_dev(c, m, n, nnz, row_ptr, col_ind, exact, descr, false);
}

extern "C" int lal_csrmv_descr_destroy(lal_csrmv_descr d)
{
    if(!d)
        return LAL_OK;
    (void)hipFree(d->d_blocks);
    (void)hipFree(d->d_c16);
    (void)hipFree(d->d_cbase);
    delete d;
    return LAL_OK;
}

extern "C" int lal_csrmv_f64(lal_handle c, lal_csrmv_descr d, int m, int n, int nnz,
                             double alpha, const int* rp, const int* ci, const double* v,
                             const double* x, double beta, double* y)
{
    return lal::csrmv_f64(c, d, m, n, nnz, alpha, rp, ci, v, x, beta, y);
}

extern "C" int lal_csrmv_f32(lal_handle c, lal_csrmv_descr d, int m, int n, int nnz,
                             float alpha, const int* rp, const int* ci, const float* v,
                             const float* x, float beta, float* y)
{
    return lal::csrmv_f32(c, d, m, n, nnz, alpha, rp, ci, v, x, beta, y);
}

extern "C" int lal_residual_f64(lal_handle c, lal_csrmv_descr d, int m, int n, int nnz,
                                const int* rp, const int* ci, const double* v, const double* x,
                                const double* b, double* r)
{
    return lal::residual_f64(c, d, m, n, nnz, rp, ci, v, x, b, r);
}

extern "C" int lal_residual_f32(lal_handle c, lal_csrmv_descr d, int m, int n, int nnz,
                                const int* rp, const int* ci, const float* v, const float* x,
                                const float* b, float* r)
{
    return lal::residual_f32(c, d, m, n, nnz, rp, ci, v, x, b, r);
}

extern "C" int lal_diagonal_f64(lal_handle c, int m, const int* rp, const int* ci,
                                const double* v, double* d)
{
    LAL_REQUIRE(c != nullptr && m >= 0);
    LAL_REQUIRE(m == 0 || (rp != nullptr && d != nullptr));
    return lal::diagonal_f64(c, m, rp, ci, v, d);
}
