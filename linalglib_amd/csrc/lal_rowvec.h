// Fast (non-exact) CSR kernels: products staged in LDS.
//
// A 256-lane workgroup owns one row block of the analysis (at most 256 rows,
// at most CAP nonzeros unless a single row is longer).  Phase 1 streams the
// block's column indices and values with non-temporal loads (they are read
// once; x keeps the caches), gathers x and stores the products a_ij * x_j in
// LDS -- 8 bytes per nonzero, so more workgroups fit on a CU than when value
// and x are staged separately.  Phase 2 sums each row's products: one lane per
// row in dense blocks, a group of G lanes with a fixed butterfly in sparse
// ones.  The sum order differs from the host's sequential fma, so results
// agree to rounding, deterministically; the bit-exact path is the row-block
// kernel of lal_csrmv.hip.
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
    };

    template <typename T>
    int rowvec_launch(lal_context* c, int mode, const rv_args& a);
}
