// Fast (non-exact) CSR kernels.
//
// A 256-lane workgroup owns one row block of the analysis (at most 256 rows,
// at most 2048 nonzeros unless a single row is longer).  The main kernel
// (16-byte aligned column and value arrays) copies the block's column indices
// and values into LDS with 16-byte non-temporal loads, four nonzeros per lane,
// and then lets every row -- one lane, or a group of G lanes with a fixed
// butterfly in blocks of few long rows -- walk its own nonzeros in LDS and
// gather x itself, eight (G = 1) or four gathers in flight per lane.  A gather
// instruction then touches the same stencil offset of consecutive rows.  For
// the rectangular transfer operators (P: 4-7 nonzeros per row, R: their
// transposes) the lanes instead gather x for the consecutive nonzeros they
// loaded and stage the products, which wastes no lanes on short rows.
// Unaligned arrays take the product-staging kernel: lanes gather x for
// consecutive nonzeros and stage a_ij * x_j in LDS before the row sums.
// Column indices and values are read once per call, so they are streamed
// (non-temporal) and x keeps the caches -- except for an operator marked
// `keep` that fits lal::CACHE_KEEP_BYTES (a hierarchy's finest level, read
// several times per solver iteration), whose ordinary loads let it stay in
// the Infinity Cache between calls.  The sum order differs from the
// host's sequential fma, so results agree to rounding, deterministically; the
// bit-exact path is the row-block kernel of lal_csrmv.hip.
//
// All level operations are one kernel with different gathers and epilogues:
//   SPMV      y = alpha*(A x) + beta*y
//   RESIDUAL  y = b - A x
//   JACOBI    y = x + (b - A x) * dinv     (= (b - sum_{k!=i} a_ik x_k)/a_ii)
//   JACOBI0   y = b * dinv and r = b - A (b * dinv)  (one sweep from x = 0,
//                                           fused with the residual)
//   RESTRICT  y = A x and r = y * dinv     (restriction that also writes the
//                                           coarse level's first sweep)
// dinv holds 1/a_ii, so no division is left in the inner loops.
#pragma once

#include "lal_spmv.h"

namespace lal
{
    enum
    {
        RV_SPMV     = 0,
        RV_RESIDUAL = 1,
        RV_JACOBI   = 2,
        RV_JACOBI0  = 3,
        RV_RESTRICT = 4 // y = A x and r = y * dinv: the coarse right-hand side
                        // and the first Jacobi iterate from zero on it
    };

    struct rv_args
    {
        int         m;
        int         nnz;
        int         by_nnz; // gather over consecutive nonzeros (transfer operators)
        const int*  blocks; // analysis row blocks (nb + 1)
        int         nb;
        const int*  rp;
        const int*  ci;
        const void* val;
        const void* x;
        const void* d;
        const void* b;
        void*       y;
        void*       r;
        double      alpha;
        double      beta;
        int         win; // LDS window in entries (lal_csrmv_info::win)
        int         keep = 0; // the matrix may stay in the cache (lal_csrmv_info::keep)
        // 16-bit columns and their per-block bases (lal_csrmv_info::d_c16), or null
        const unsigned short* c16   = nullptr;
        const int*            cbase = nullptr;
    };

    // Builds the 16-bit column copy of a fast-path descriptor's matrix
    // (d->d_c16, d->d_cbase) when every row block's columns span at most
    // 65535; otherwise leaves the descriptor as it is.  Results of the row
    // kernels are the same either way: only the bytes read change.
    int compact_columns(lal_context* c, lal_csrmv_info* d, int nnz, const int* ci);

    // The descriptor's 16-bit columns into a's fields.
    inline void use_c16(rv_args& a, const lal_csrmv_info* d)
    {
        a.c16   = d->d_c16;
        a.cbase = d->d_cbase;
    }

    template <typename T>
    int rowvec_launch(lal_context* c, int mode, const rv_args& a);
}
