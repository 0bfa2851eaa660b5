// Single-workgroup coarse tail of the V-cycle (see lal_coarse.hip).
#pragma once

#include "lal_internal.h"

namespace lal
{
    // Device-resident description of one level handled by the coarse tail.
    // P and R are only used for levels above the coarsest.
    template <typename T>
    struct coarse_level
    {
        int        m;
        const int* rp;
        const int* ci;
        const T*   v;
        const T*   dinv;
        const int* prp;
        const int* pci;
        const T*   pv;
        const int* rrp;
        const int* rci;
        const T*   rv;
        T*         x;
        T*         t;
        T*         b;
        T*         r;
    };

    // Runs the tail for nl levels (the last one the coarsest).  Level 0 of
    // the array takes its right-hand side from its b buffer; the result of
    // that level lands in `result` when given, else in its t buffer.
    // kind 0: the coarsest is 1x1; 1: dense inverse `inv`.
    template <typename T>
    int coarse_tail(lal_context* c, const coarse_level<T>* d_levels, int nl, int kind,
                    const double* inv, T* result);

    // Levels whose operator has at most this many nonzeros go to the tail.
    constexpr int COARSE_TAIL_NNZ = 8192;
}
