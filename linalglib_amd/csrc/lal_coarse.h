// Single-workgroup coarse tail of the V-cycle (see lal_coarse.hip).
#pragma once

#include "lal_internal.h"

namespace lal
{
    // Levels the tail can hold, and the LDS it may use for them.
    constexpr int TAIL_MAX_LEVELS = 12;
    constexpr int TAIL_LDS_BYTES  = 64 * 1024;

    // Word (4-byte) offsets of one level's arrays inside the tail's LDS
    // image.  P and R are absent (-1) on the coarsest level.
    struct tail_level
    {
        int m;
        int rp, ci, v, dinv;     // A and 1/a_ii
        int prp, pci, pv;        // P (m x m_next)
        int rrp, rci, rv;        // R (m_next x m)
        int x, t, b, r;          // work vectors (LDS only, not in the image)
    };

    struct tail_desc
    {
        int        nl;        // levels in the tail, the last one the coarsest
        int        kind;      // coarsest solve: 0 the 1x1 divide, 1 dense inverse
        int        img_words; // words copied from the device image
        int        lds_words; // image plus work vectors
        int        inv;       // word offset of the dense inverse in LDS, -1: global
        tail_level L[TAIL_MAX_LEVELS];
    };

    // Runs the tail: the image (rp, ci, values, 1/a_ii, P, R of every tail
    // level, optionally the coarsest dense inverse) is copied into LDS once,
    // the first level's right-hand side comes from b_in, and the first
    // level's post-smoothed result lands in `result`.
    template <typename T>
    int coarse_tail(lal_context* c, const tail_desc& d, const int* image, const T* b_in,
                    const double* inv, T* result);
}
