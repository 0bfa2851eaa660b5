// Fast CSR level kernels (see lal_rowvec.h).
#include "lal_rowvec.h"

namespace
{
    using lal::SPMV_THREADS;

    // LDS bytes of a block's row starts (nrows + 1 <= SPMV_THREADS + 1 ints,
    // rounded up to 16 bytes so that the staged window stays aligned).
    constexpr int RP_LDS_BYTES = (int)(((SPMV_THREADS + 1) * sizeof(int) + 15) & ~15);

    template <typename T>
    __device__ __forceinline__ T stream_load(const T* p)
    {
        return __builtin_nontemporal_load(p);
    }

    typedef int            int4v __attribute__((ext_vector_type(4)));
    typedef unsigned short us4v __attribute__((ext_vector_type(4)));
    typedef float  flt4v __attribute__((ext_vector_type(4)));
    typedef double dbl2v __attribute__((ext_vector_type(2)));

    // Four consecutive matrix values read with 16-byte non-temporal loads.
    template <typename T>
    struct quad;
    template <>
    struct quad<float>
    {
        flt4v v;
        template <bool KEEP = false>
        __device__ __forceinline__ void load(const float* p)
        {
            if(KEEP)
                v = *reinterpret_cast<const flt4v*>(p);
            else
                v = __builtin_nontemporal_load(reinterpret_cast<const flt4v*>(p));
        }
        __device__ __forceinline__ void store(float* s) const
        {
            *reinterpret_cast<flt4v*>(s) = v;
        }
    };
    template <>
    struct quad<double>
    {
        dbl2v a, b;
        template <bool KEEP = false>
        __device__ __forceinline__ void load(const double* p)
        {
            if(KEEP)
            {
                a = reinterpret_cast<const dbl2v*>(p)[0];
                b = reinterpret_cast<const dbl2v*>(p)[1];
            }
            else
            {
                a = __builtin_nontemporal_load(reinterpret_cast<const dbl2v*>(p));
                b = __builtin_nontemporal_load(reinterpret_cast<const dbl2v*>(p) + 1);
            }
        }
        __device__ __forceinline__ void store(double* s) const
        {
            reinterpret_cast<dbl2v*>(s)[0] = a;
            reinterpret_cast<dbl2v*>(s)[1] = b;
        }
    };

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

    // y (and r) of one row from its sum acc = (A u)_row, per operation.
    template <typename T, int MODE>
    __device__ __forceinline__ void row_epilogue(int row, T acc, const T* __restrict__ x,
                                                 const T* __restrict__ d,
                                                 const T* __restrict__ b, T alpha, T beta,
                                                 T* __restrict__ y, T* __restrict__ r)
    {
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

    // Row sum over the staged nonzeros [j, e) (stride g): UR column indices
    // and values come from LDS, then UR gathers are in flight at once.
    template <typename T, int MODE, int UR>
    __device__ __forceinline__ T staged_row_sum(const int* s_c, const T* s_v, int j, int e, int g,
                                                const T* __restrict__ x,
                                                const T* __restrict__ d,
                                                const T* __restrict__ b)
    {
        T acc = 0;
        if(j >= e)
            return acc;
        const int csafe = s_c[j];
        for(; j < e; j += UR * g)
        {
            int  cc[UR];
            T    vv[UR], xx[UR];
            bool ok[UR];
#pragma unroll
            for(int u = 0; u < UR; u++)
            {
                ok[u] = j + u * g < e;
                cc[u] = ok[u] ? s_c[j + u * g] : csafe;
                vv[u] = ok[u] ? s_v[j + u * g] : (T)0;
            }
#pragma unroll
            for(int u = 0; u < UR; u++)
                xx[u] = gather<T, MODE>(x, d, b, cc[u]);
#pragma unroll
            for(int u = 0; u < UR; u++)
            {
                const T p = vv[u] * xx[u];
                acc += ok[u] ? p : (T)0;
            }
        }
        return acc;
    }

    // Four consecutive column indices at q (a multiple of 4): 16 bytes of
    // int32, or 8 bytes of 16-bit offsets from the block's base cb.
    template <bool KEEP, bool C16>
    __device__ __forceinline__ int4v load_cols(const int* __restrict__ ci,
                                               const unsigned short* __restrict__ c16, int cb,
                                               int q)
    {
        if(C16)
        {
            const us4v* p = reinterpret_cast<const us4v*>(c16 + q);
            const us4v  w = KEEP ? *p : __builtin_nontemporal_load(p);
            return int4v{cb + (int)w.x, cb + (int)w.y, cb + (int)w.z, cb + (int)w.w};
        }
        const int4v* p = reinterpret_cast<const int4v*>(ci + q);
        return KEEP ? *p : __builtin_nontemporal_load(p);
    }

    // Quad-staged row-gather kernel (16-byte aligned ci and val).
    //
    // Phase 1 copies the block's column indices and values into LDS with
    // 16-byte non-temporal loads, four consecutive nonzeros per lane, over a
    // window aligned down to a multiple of four (WIN, a multiple of four, is
    // the widest such window of the matrix's blocks and at most CAP + 4
    // entries, so a block of at most CAP nonzeros always fits one window).  Phase 2 lets each
    // row (one lane, or G lanes for short blocks) walk its own nonzeros in LDS
    // and gather x itself: one gather instruction then covers the same
    // stencil offset of 64 consecutive rows, a few cache lines, instead of
    // every offset of a handful of rows.  Fewer and wider memory instructions
    // and compact gathers are what lift this kernel over the product-staging
    // one on CDNA.
    template <typename T, int MODE, bool KEEP, bool C16>
    __global__ void __launch_bounds__(SPMV_THREADS)
        quad_kernel(int m, int nnz, const int* __restrict__ blocks, int nb, int win,
                    const int* __restrict__ rp, const int* __restrict__ ci,
                    const unsigned short* __restrict__ c16, const int* __restrict__ cbase,
                    const T* __restrict__ val, const T* __restrict__ x,
                    const T* __restrict__ d, const T* __restrict__ b, T alpha, T beta,
                    T* __restrict__ y, T* __restrict__ r)
    {
        // the block's row starts, then win entries of columns, then win
        // values (win: a multiple of 4)
        extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
        int* s_rp = reinterpret_cast<int*>(s_raw);
        int* s_c  = reinterpret_cast<int*>(s_raw + RP_LDS_BYTES);
        T*   s_v  = reinterpret_cast<T*>(s_raw + RP_LDS_BYTES + sizeof(int) * win);
        const int WIN = win;

        const int   tid  = threadIdx.x;
        const int   tile = lal::xcd_tile(blockIdx.x, nb);
        const int4v mt   = reinterpret_cast<const int4v*>(blocks + lal::block_meta_offset(nb))[tile];
        const int   r0 = mt.x, r1 = mt.y;
        const int  nrows  = r1 - r0;
        const int  g      = lal::spmv_group(nrows, 0);
        const int  lrow   = tid / g;
        const int  lig    = tid & (g - 1);
        const bool active = lrow < nrows;
        const int  row    = r0 + lrow;
        const int  base   = mt.z;
        const int  end    = mt.w;
        const int  nnz4   = nnz & ~3;
        const int  cb     = C16 ? cbase[tile] : 0; // this block's 16-bit column base
        // The row bounds travel with the first window's loads (through LDS)
        // instead of being awaited before them: rp[r0] and rp[r1] are the
        // descriptor's base and end, lane i loads rp[r0 + i] for 0 < i < nrows.
        const int rpv = (tid > 0 && tid < nrows) ? rp[r0 + tid] : 0;
        int       rs = 0, re = 0;
        bool      bounds = false;

        T acc = 0;
        for(int ws = base & ~3; ws < end; ws += WIN)
        {
            const int we   = min(ws + WIN, end);
            const int vend = min((we + 3) & ~3, nnz4); // quads inside the arrays
            for(int q0 = ws + 4 * tid; q0 < vend; q0 += 8 * SPMV_THREADS)
            {
                int4v   cq[2];
                quad<T> vq[2];
                int     qq[2];
#pragma unroll
                for(int u = 0; u < 2; u++)
                {
                    qq[u] = min(q0 + 4 * SPMV_THREADS * u, vend - 4);
                    cq[u] = load_cols<KEEP, C16>(ci, c16, cb, qq[u]);
                    vq[u].template load<KEEP>(val + qq[u]);
                }
                // Every quad is stored unconditionally: a lane whose second
                // quad lies past the window's end stores the window's last
                // quad again, the same values its owner stores.  A stored
                // value used only under a condition lets the compiler sink
                // its load into that branch, after the first quad's wait,
                // which makes two HBM round trips of one.
#pragma unroll
                for(int u = 0; u < 2; u++)
                {
                    *reinterpret_cast<int4v*>(s_c + (qq[u] - ws)) = cq[u];
                    vq[u].store(s_v + (qq[u] - ws));
                }
            }
            // the arrays' last nnz % 4 entries
            for(int k = max(vend, ws) + tid; k < we; k += SPMV_THREADS)
            {
                s_c[k - ws] = C16 ? cb + (int)c16[k] : ci[k];
                s_v[k - ws] = val[k];
            }
            if(!bounds)
            {
                if(tid > 0 && tid < nrows)
                    s_rp[tid] = rpv;
                if(tid == 0)
                {
                    s_rp[0]     = base;
                    s_rp[nrows] = end;
                }
            }
            __syncthreads();
            if(!bounds)
            {
                rs     = active ? s_rp[lrow] : 0;
                re     = active ? s_rp[lrow + 1] : 0;
                bounds = true;
            }
            if(active)
            {
                const int a = max(rs, ws) - ws + lig;
                const int e = min(re, we) - ws;
                acc += g == 1 ? staged_row_sum<T, MODE, 8>(s_c, s_v, a, e, 1, x, d, b)
                              : staged_row_sum<T, MODE, 4>(s_c, s_v, a, e, g, x, d, b);
            }
            __syncthreads();
        }
        if(g > 1)
            acc = lal::group_sum(acc, g);
        if(!active || lig != 0)
            return;
        row_epilogue<T, MODE>(row, acc, x, d, b, alpha, beta, y, r);
    }

    // Quad-staged kernel that gathers by nonzero (16-byte aligned ci and val):
    // each lane gathers x for the four consecutive nonzeros it loaded and
    // stores the products in LDS; the rows then sum their products.  Used for
    // the rectangular transfer operators, whose short and uneven rows would
    // leave lanes of the per-row gather idle.
    template <typename T, int MODE, bool C16>
    __global__ void __launch_bounds__(SPMV_THREADS)
        quad_nnz_kernel(int m, int nnz, const int* __restrict__ blocks, int nb, int win,
                        const int* __restrict__ rp, const int* __restrict__ ci,
                        const unsigned short* __restrict__ c16, const int* __restrict__ cbase,
                        const T* __restrict__ val, const T* __restrict__ x,
                        const T* __restrict__ d, const T* __restrict__ b, T alpha, T beta,
                        T* __restrict__ y, T* __restrict__ r)
    {
        // the block's row starts, then win products
        extern __shared__ __attribute__((aligned(16))) unsigned char s_raw[];
        int*      s_rp = reinterpret_cast<int*>(s_raw);
        T*        s_p  = reinterpret_cast<T*>(s_raw + RP_LDS_BYTES);
        const int WIN  = win;

        const int   tid  = threadIdx.x;
        const int   tile = lal::xcd_tile(blockIdx.x, nb);
        const int4v mt   = reinterpret_cast<const int4v*>(blocks + lal::block_meta_offset(nb))[tile];
        const int   r0 = mt.x, r1 = mt.y;
        const int  nrows  = r1 - r0;
        const int  g      = lal::spmv_group(nrows, 0);
        const int  lrow   = tid / g;
        const int  lig    = tid & (g - 1);
        const bool active = lrow < nrows;
        const int  row    = r0 + lrow;
        const int  base   = mt.z;
        const int  end    = mt.w;
        const int  nnz4   = nnz & ~3;
        const int  cb     = C16 ? cbase[tile] : 0; // this block's 16-bit column base
        // row bounds through LDS, loaded with the first window (quad_kernel)
        const int rpv = (tid > 0 && tid < nrows) ? rp[r0 + tid] : 0;
        int       rs = 0, re = 0;
        bool      bounds = false;

        T acc = 0;
        for(int ws = base & ~3; ws < end; ws += WIN)
        {
            const int we   = min(ws + WIN, end);
            const int vend = min((we + 3) & ~3, nnz4);
            for(int q0 = ws + 4 * tid; q0 < vend; q0 += 8 * SPMV_THREADS)
            {
                // two quads per lane, both loaded before either's gathers (a
                // second quad past the window's end repeats its last quad and
                // stores the same products as that quad's owner)
                int4v   cq[2];
                quad<T> vq[2];
                int     qq[2];
#pragma unroll
                for(int u = 0; u < 2; u++)
                {
                    qq[u] = min(q0 + 4 * SPMV_THREADS * u, vend - 4);
                    cq[u] = load_cols<false, C16>(ci, c16, cb, qq[u]);
                    vq[u].load(val + qq[u]);
                    // a 16-bit column of a neighbouring block's entry (a quad
                    // straddling the block's ends) decodes against this
                    // block's base: gather at the base instead (its product
                    // belongs to no row of this block)
                    if(C16)
#pragma unroll
                        for(int e = 0; e < 4; e++)
                            if(qq[u] + e < base || qq[u] + e >= end)
                                cq[u][e] = cb;
                }
                T v8[8], x8[8];
#pragma unroll
                for(int u = 0; u < 2; u++)
                    vq[u].store(v8 + 4 * u);
#pragma unroll
                for(int u = 0; u < 2; u++)
#pragma unroll
                    for(int e = 0; e < 4; e++)
                        x8[4 * u + e] = gather<T, MODE>(x, d, b, cq[u][e]);
#pragma unroll
                for(int u = 0; u < 2; u++)
#pragma unroll
                    for(int e = 0; e < 4; e++)
                        s_p[qq[u] - ws + e] = v8[4 * u + e] * x8[4 * u + e];
            }
            for(int k = max(vend, ws) + tid; k < we; k += SPMV_THREADS)
                s_p[k - ws] = val[k] * gather<T, MODE>(x, d, b,
                                                       C16 ? (k < base ? cb : cb + (int)c16[k])
                                                           : ci[k]);
            if(!bounds)
            {
                if(tid > 0 && tid < nrows)
                    s_rp[tid] = rpv;
                if(tid == 0)
                {
                    s_rp[0]     = base;
                    s_rp[nrows] = end;
                }
            }
            __syncthreads();
            if(!bounds)
            {
                rs     = active ? s_rp[lrow] : 0;
                re     = active ? s_rp[lrow + 1] : 0;
                bounds = true;
            }
            if(active)
            {
                const int a = max(rs, ws) - ws + lig;
                const int e = min(re, we) - ws;
                for(int j = a; j < e; j += g)
                    acc += s_p[j];
            }
            __syncthreads();
        }
        if(g > 1)
            acc = lal::group_sum(acc, g);
        if(!active || lig != 0)
            return;
        row_epilogue<T, MODE>(row, acc, x, d, b, alpha, beta, y, r);
    }

    template <typename T, int MODE, bool C16>
    void launch_staged(lal_context* c, const lal::rv_args& a, int nb)
    {
        // LDS sized to the widest block window of this matrix, so that short
        // blocks leave room for more resident workgroups
        const int    win  = a.win;
        const size_t cbyt = C16 ? sizeof(unsigned short) : sizeof(int);
        if(a.by_nnz)
            quad_nnz_kernel<T, MODE, C16><<<nb, SPMV_THREADS, RP_LDS_BYTES + sizeof(T) * win,
                                           c->stream>>>(
                a.m, a.nnz, a.blocks, nb, win, a.rp, a.ci, a.c16, a.cbase, (const T*)a.val,
                (const T*)a.x, (const T*)a.d, (const T*)a.b, (T)a.alpha, (T)a.beta, (T*)a.y,
                (T*)a.r);
        else if(a.keep && (sizeof(T) + cbyt) * (size_t)a.nnz <= lal::CACHE_KEEP_BYTES)
            quad_kernel<T, MODE, true, C16><<<nb, SPMV_THREADS,
                                              RP_LDS_BYTES + (sizeof(int) + sizeof(T)) * win,
                                              c->stream>>>(
                a.m, a.nnz, a.blocks, nb, win, a.rp, a.ci, a.c16, a.cbase, (const T*)a.val,
                (const T*)a.x, (const T*)a.d, (const T*)a.b, (T)a.alpha, (T)a.beta, (T*)a.y,
                (T*)a.r);
        else
            quad_kernel<T, MODE, false, C16><<<nb, SPMV_THREADS,
                                               RP_LDS_BYTES + (sizeof(int) + sizeof(T)) * win,
                                               c->stream>>>(
                a.m, a.nnz, a.blocks, nb, win, a.rp, a.ci, a.c16, a.cbase, (const T*)a.val,
                (const T*)a.x, (const T*)a.d, (const T*)a.b, (T)a.alpha, (T)a.beta, (T*)a.y,
                (T*)a.r);
    }

    template <typename T, int MODE>
    int launch(lal_context* c, const lal::rv_args& a)
    {
        const int nb = a.blocks ? a.nb : (a.m + SPMV_THREADS - 1) / SPMV_THREADS;
        if(nb == 0)
            return LAL_OK;
        const bool val16 = (reinterpret_cast<uintptr_t>(a.val) & 15) == 0;
        if(a.blocks && val16 && a.c16 && a.cbase)
            launch_staged<T, MODE, true>(c, a, nb);
        else if(a.blocks && val16 && (reinterpret_cast<uintptr_t>(a.ci) & 15) == 0)
            launch_staged<T, MODE, false>(c, a, nb);
        else
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
