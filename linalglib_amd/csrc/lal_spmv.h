// Row-block CSR kernels shared by SpMV, residual and the smoothers.
//
// A workgroup of 256 lanes owns a contiguous block of rows [r0, r1) chosen by
// the analysis (at most 256 rows and, unless a single row is longer, at most
// CAP nonzeros).  The block's nonzero range is staged through LDS in chunks of
// CAP entries with coalesced loads: the matrix value and the gathered x value
// of every nonzero land side by side, so the row sums read LDS only.  With
// G = 1 each row is summed by one lane in CSR order, which reproduces the
// host loop bit for bit; sparse blocks (few, long rows) split each row over a
// group of G lanes and combine the partial sums with a fixed butterfly.
#pragma once

#include "lal_internal.h"

struct lal_csrmv_info
{
    int  m        = 0;
    int  nnz      = 0;
    int  nblocks  = 0;
    int  exact    = 0;       // 1: sequential row sums (bit-identical to the host)
    int  by_nnz   = 0;       // fast path: lanes gather x over consecutive nonzeros
                             // (rectangular transfer operators) instead of per row
    int  win      = 4;       // widest 4-aligned nonzero window of a block (<= 2052)
    int  keep     = 0;       // fast path: the matrix may be read with
                             // cache-allocating loads instead of streaming
                             // ones when it fits lal::CACHE_KEEP_BYTES
    int* d_blocks = nullptr; // nblocks + 1 row offsets, then per block
                             // {r0, r1, row_ptr[r0], row_ptr[r1]} at
                             // lal::block_meta_offset(nblocks)
    // 16-bit columns (fast path of the operators a hierarchy owns, set by
    // lal::compact_columns): entry k of block b stores col_ind[k] - d_cbase[b];
    // null when not built or a block's columns span more than 65535
    unsigned short* d_c16   = nullptr;
    int*            d_cbase = nullptr;
};

namespace lal
{
    // Position (in ints, a multiple of 4) of the per-block quadruples in the
    // descriptor's block array.
    __host__ __device__ inline size_t block_meta_offset(int nblocks)
    {
        return ((size_t)nblocks + 1 + 3) & ~(size_t)3;
    }

    constexpr int SPMV_THREADS = 256;

    // Largest matrix (values and column indices at the precision of the call)
    // that a descriptor marked `keep` reads with cache-allocating loads: three
    // quarters of MI355X's 256 MB Infinity Cache, leaving room for the vectors.
    constexpr size_t CACHE_KEEP_BYTES = (size_t)192 << 20;

    template <typename T>
    struct spmv_cap
    {
        static constexpr int value = 2048;
    };

    // Lanes per row for a block of `rows` rows.
    __device__ __forceinline__ int spmv_group(int rows, int exact)
    {
        if(exact || rows >= 128)
            return 1;
        int g = 1;
        while(g < 64 && 2 * g * rows <= SPMV_THREADS)
            g <<= 1;
        return g;
    }

    template <typename T>
    __device__ __forceinline__ T group_sum(T v, int g)
    {
        for(int o = g >> 1; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        return v;
    }

    __device__ __forceinline__ void block_rows(const int* __restrict__ blocks, int m, int tile,
                                               int* r0, int* r1)
    {
        if(blocks)
        {
            *r0 = blocks[tile];
            *r1 = blocks[tile + 1];
        }
        else
        {
            *r0 = tile * SPMV_THREADS;
            *r1 = min(m, *r0 + SPMV_THREADS);
        }
    }

    inline int spmv_grid(const lal_csrmv_info* d, int m)
    {
        return d ? d->nblocks : (m + SPMV_THREADS - 1) / SPMV_THREADS;
    }
}
