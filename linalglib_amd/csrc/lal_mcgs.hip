// Multicolour Gauss-Seidel smoothers for the AMG cycles (smoother 5: one
// forward colour sweep; 6: forward then backward, a symmetric sweep).
//
// LinAlgLib's Gauss-Seidel (iterative_solvers/classic/gauss_seidel.cpp:50-87,
// applied at amg.cpp:127-132) updates rows in index order, a sequential chain
// that the level-scheduled sweep of lal_cycles.hip reproduces bit for bit.
// A multicolour sweep is a different smoother with the same update formula,
// x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii: the rows are coloured so that no
// two rows of one colour are coupled (a_ij != 0 or a_ji != 0), and the colours
// are swept in order, every row of a colour at once.  Within a colour no row
// reads another's new value, so a colour is one parallel pass; across colours
// each row reads the newest values of the colours before it.  The forward and
// backward sweeps are each other's transpose, so a V-cycle with symmetric
// sweeps (6) before and after the coarse correction is a symmetric
// preconditioner for CG.
//
// The colouring is greedy first fit (on the host, on first use of a level):
// each row takes the smallest colour that no already coloured row coupled to
// it holds, rows taken in index order or, when that needs fewer colours, in
// smallest-last order (colour_rows).  On a 7-point grid operator in
// lexicographic order it is the red-black colouring.  The level keeps a copy of its matrix with the
// rows grouped by colour (off-diagonal entries only, and 1/a_ii per row), so
// a colour's rows are one contiguous stream.  Colours of many rows are one
// launch each, with 1 to 64 lanes per row by the level's mean row length;
// levels small enough to keep x and b in one workgroup's LDS sweep every
// colour in one launch of one workgroup (mc_wg_kernel).
#include "lal_hier.h"

#include <vector>

namespace
{
    using lal::csr_dev;
    using lal::mc_sched;

    constexpr int MC_WG_U       = 4;          // entries per lane and chunk
    constexpr int MC_WG_PAD     = 64;         // doubles past the product buffers
    constexpr int MC_WG_PF      = 3;          // chunks whose loads are issued ahead, plus one
    constexpr int MC_WG_LDS_MAX = 160 * 1024; // LDS of one workgroup on gfx950
    constexpr int MC_WG_MAX_NNZ = 262144;     // entries of a one-workgroup level

    typedef int int4v __attribute__((ext_vector_type(4)));

    // Sum over aligned groups of 2^lg <= 16 lanes, left in every lane of the
    // group: quad permutations (lane ^ 1, lane ^ 2), then the half-row and
    // row mirrors, each adding the other half's partial sum.
    __device__ __forceinline__ double group_sum16(double v, int lg)
    {
        if(lg >= 1)
            v = lal::dpp_add<0xB1>(v); // quad_perm [1, 0, 3, 2]
        if(lg >= 2)
            v = lal::dpp_add<0x4E>(v); // quad_perm [2, 3, 0, 1]
        if(lg >= 3)
            v = lal::dpp_add<0x141>(v); // row_half_mirror
        if(lg >= 4)
            v = lal::dpp_add<0x140>(v); // row_mirror
        return v;
    }

    // Rows t in [0, cnt) of one colour, LANES lanes per row.  A row's entries
    // are taken 4 * LANES at a time, four per lane, all loads of a round
    // issued before the first is used (an index past the row reads the row's
    // last entry and contributes zero), so a row of up to 4 * LANES entries
    // costs one round of dependent loads.
    template <int LANES>
    __global__ void __launch_bounds__(256)
        mc_colour_kernel(int cnt, const int* __restrict__ perm, const int* __restrict__ prp,
                         const int* __restrict__ pci, const double* __restrict__ pv,
                         const double* __restrict__ pdinv, const double* __restrict__ b,
                         double* x)
    {
        const int gt  = blockIdx.x * 256 + threadIdx.x;
        const int t   = gt / LANES;
        const int lig = gt % LANES;
        if(t >= cnt) // whole lane groups leave together (LANES divides 64)
            return;
        const int rs = prp[t], re = prp[t + 1];
        // the row's position, b_i and 1/a_ii do not depend on the sum: load
        // them now, beside the row bounds, rather than after the reduction
        const int    i  = perm[t];
        const double di = pdinv[t];
        const double bi = b[i];
        double       s  = 0.0;
        for(int k0 = rs + lig; k0 - lig < re; k0 += 4 * LANES) // re > rs inside
        {
            int    c[4];
            double v[4], xx[4];
#pragma unroll
            for(int u = 0; u < 4; u++)
            {
                const int k = min(k0 + u * LANES, re - 1);
                c[u]        = pci[k];
                v[u]        = pv[k];
            }
#pragma unroll
            for(int u = 0; u < 4; u++)
                xx[u] = x[c[u]];
#pragma unroll
            for(int u = 0; u < 4; u++)
                s += k0 + u * LANES < re ? v[u] * xx[u] : 0.0;
        }
        s = lal::lane_sum<LANES>(s);
        if(lig == 0)
            x[i] = (bi - s) * di;
    }

    // Buffer resource over `bytes` bytes at p (gfx9-family raw buffer): a
    // load at a byte offset past the end returns 0 and touches no memory.
    __device__ __forceinline__ __amdgpu_buffer_rsrc_t buf_rsrc(const void* p, int bytes)
    {
        return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), 0, bytes, 0x00020000);
    }

    __device__ __forceinline__ int buf_i32(__amdgpu_buffer_rsrc_t r, int i)
    {
        return __builtin_amdgcn_raw_buffer_load_b32(r, i * 4, 0, 0);
    }

    __device__ __forceinline__ double buf_f64(__amdgpu_buffer_rsrc_t r, int i)
    {
        return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, i * 8, 0, 0));
    }

    // Every colour of a small level in one workgroup of NT lanes, with x, b
    // and the chunk plan in LDS (mc_wg_lds_bytes).  The host cuts each
    // colour's rows, in order, into chunks of whole rows with at most
    // NT * MC_WG_U entries and at most NT / G rows, G = 2^lg <= 16 lanes per
    // row; the plan holds two int4 per chunk ({first entry, end entry, first
    // row, end row} and {1 when the chunk ends its colour}), the forward
    // sweep's chunks, then the backward sweep's, each padded with empty
    // chunks to a multiple of MC_WG_PF.  Per chunk the lanes multiply MC_WG_U
    // entries each by x read from LDS into one of two LDS product buffers,
    // and after a barrier the G lanes of row slot k add the products of the
    // chunk's row k (lane-strided, then group_sum16) and the first of them
    // writes the row's new x.  Rows of one colour are uncoupled, so the next
    // chunk of the same colour may read x while these rows are written; a new
    // colour waits at a second barrier.  Nothing but x depends on an earlier
    // chunk, so the loads of the next MC_WG_PF - 1 chunks (entries and the
    // rows' bounds, positions and 1/a_ii) are in flight while a chunk is
    // worked on; the barriers order LDS only and leave them outstanding.  The
    // matrix is read through buffer resources, so an entry or row index past
    // its array (a chunk's last lanes, slots past its rows) loads 0 without
    // an address clamp; such values are never used.
    template <int NT>
    __global__ void __launch_bounds__(NT)
        mc_wg_kernel(int m, int nnz, int nsteps, int lg, const int4v* __restrict__ plan,
                     const int* __restrict__ perm, const int* __restrict__ prp,
                     const int* __restrict__ pci, const double* __restrict__ pv,
                     const double* __restrict__ pdinv, const double* __restrict__ b, double* x)
    {
        constexpr int CH = NT * MC_WG_U;
        constexpr int PF = MC_WG_PF;
        extern __shared__ double s_mem[];
        double*   s_x  = s_mem;
        double*   s_b  = s_mem + m;
        double*   s_p  = s_mem + 2 * m; // two buffers of CH products
        int4v*    s_pl = reinterpret_cast<int4v*>(s_p + 2 * CH + MC_WG_PAD); // the plan
        const int tid  = threadIdx.x;
        const int G    = 1 << lg;
        const int S    = NT >> lg; // row slots
        const int slot = tid >> lg;
        const int lig  = tid & (G - 1);
        const auto r_ci = buf_rsrc(pci, nnz * 4);
        const auto r_v  = buf_rsrc(pv, nnz * 8);
        const auto r_rp = buf_rsrc(prp, (m + 1) * 4);
        const auto r_pm = buf_rsrc(perm, m * 4);
        const auto r_dv = buf_rsrc(pdinv, m * 8);
        for(int i = tid; i < m; i += NT)
            s_x[i] = x[i];
        for(int i = tid; i < m; i += NT)
            s_b[i] = b[perm[i]];
        for(int i = tid; i < 2 * nsteps; i += NT)
            s_pl[i] = plan[i];
        lal::lds_barrier();

        int    C[PF][MC_WG_U], RS[PF], RE[PF], RR[PF];
        double V[PF][MC_WG_U], DV[PF];

        // chunk j's loads into register set r
#define MC_WG_FETCH(j, r)                                                                    \
    do                                                                                       \
    {                                                                                        \
        const int4v h_  = s_pl[2 * (j)];                                                     \
        const int   c0_ = __builtin_amdgcn_readfirstlane(h_.x);                              \
        const int   t_  = __builtin_amdgcn_readfirstlane(h_.z) + slot;                       \
        _Pragma("unroll") for(int u = 0; u < MC_WG_U; u++)                                   \
        {                                                                                    \
            C[r][u] = buf_i32(r_ci, c0_ + tid + NT * u);                                     \
            V[r][u] = buf_f64(r_v, c0_ + tid + NT * u);                                      \
        }                                                                                    \
        RS[r] = buf_i32(r_rp, t_);                                                           \
        RE[r] = buf_i32(r_rp, t_ + 1);                                                       \
        RR[r] = buf_i32(r_pm, t_);                                                           \
        DV[r] = buf_f64(r_dv, t_);                                                           \
    } while(0)

        // chunk j from register set r: products, row sums, x of its rows
#define MC_WG_WORK(j, r)                                                                     \
    do                                                                                       \
    {                                                                                        \
        const int4v h_  = s_pl[2 * (j)];                                                     \
        const int   c0_ = __builtin_amdgcn_readfirstlane(h_.x);                              \
        const int   c1_ = __builtin_amdgcn_readfirstlane(h_.y);                              \
        const int   ta_ = __builtin_amdgcn_readfirstlane(h_.z);                              \
        const int   tb_ = __builtin_amdgcn_readfirstlane(h_.w);                              \
        const int   fl_ = __builtin_amdgcn_readfirstlane(s_pl[2 * (j) + 1].x);               \
        double*     bp_ = s_p + ((j) & 1) * CH;                                              \
        double      xv_[MC_WG_U];                                                            \
        _Pragma("unroll") for(int u = 0; u < MC_WG_U; u++) xv_[u] = s_x[C[r][u]];            \
        _Pragma("unroll") for(int u = 0; u < MC_WG_U; u++)                                   \
        {                                                                                    \
            const int e_ = tid + NT * u;                                                     \
            if(e_ < c1_ - c0_)                                                               \
                bp_[e_] = V[r][u] * xv_[u];                                                  \
        }                                                                                    \
        lal::lds_barrier();                                                                  \
        const bool ok_ = ta_ + slot < tb_;                                                   \
        const int  n_  = ok_ ? RE[r] - RS[r] : 0;                                            \
        const int  lo_ = RS[r] - c0_;                                                        \
        double     a_  = 0.0;                                                                \
        for(int k = lig; k < n_; k += 4 * G)                                                 \
        {                                                                                    \
            double p_[4];                                                                    \
            _Pragma("unroll") for(int u = 0; u < 4; u++) p_[u] = bp_[lo_ + k + u * G];       \
            _Pragma("unroll") for(int u = 0; u < 4; u++) a_ += k + u * G < n_ ? p_[u] : 0.0; \
        }                                                                                    \
        a_ = group_sum16(a_, lg);                                                            \
        if(ok_ && lig == 0)                                                                  \
            s_x[RR[r]] = (s_b[ta_ + slot] - a_) * DV[r];                                     \
        if(fl_)                                                                              \
            lal::lds_barrier();                                                              \
    } while(0)

        // Prologue: set 0 is complete before set 1 is issued (the empty asm
        // takes its values and orders memory), so that on entry to the loop,
        // as after each round of it, every set but the newest is older than
        // the newest set's loads and the waits in the loop keep MC_WG_PF - 1
        // chunks in flight.
        MC_WG_FETCH(0, 0);
        {
            int    dep = RS[0] ^ RE[0] ^ RR[0];
            double dd  = DV[0];
#pragma unroll
            for(int u = 0; u < MC_WG_U; u++)
            {
                dep ^= C[0][u];
                dd += V[0][u];
            }
            asm volatile("" ::"v"(dep), "v"(dd) : "memory");
        }
#pragma unroll
        for(int r = 1; r < PF - 1; r++)
            MC_WG_FETCH(min(r, nsteps - 1), r);
        for(int j = 0; j < nsteps; j += PF)
        {
#pragma unroll
            for(int r = 0; r < PF; r++)
            {
                MC_WG_FETCH(min(j + r + PF - 1, nsteps - 1), (r + PF - 1) % PF);
                MC_WG_WORK(j + r, r);
            }
        }
#undef MC_WG_FETCH
#undef MC_WG_WORK
        for(int i = tid; i < m; i += NT)
            x[i] = s_x[i];
    }

    // LDS of a one-workgroup sweep of NT lanes over m rows and nsteps chunks:
    // x, b, two product buffers (and MC_WG_PAD doubles that the row sums may
    // read, never use, past the second) and the plan.
    long long mc_wg_lds_bytes(int nt, int m, int nsteps)
    {
        return (2LL * m + 2LL * nt * MC_WG_U + MC_WG_PAD) * (long long)sizeof(double) +
               32LL * nsteps;
    }

    // Enables MC_WG_LDS_MAX bytes of LDS (beyond the 64 KiB default) for the
    // three one-workgroup sweeps once; false when the device does not allow it.
    bool mc_wg_lds_ok()
    {
        static int ok = -1;
        if(ok < 0)
        {
            int dev = 0, v = 0;
            ok = 0;
            if(hipGetDevice(&dev) == hipSuccess &&
               hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) ==
                   hipSuccess &&
               v >= MC_WG_LDS_MAX &&
               hipFuncSetAttribute(reinterpret_cast<const void*>(&mc_wg_kernel<64>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   MC_WG_LDS_MAX) == hipSuccess &&
               hipFuncSetAttribute(reinterpret_cast<const void*>(&mc_wg_kernel<256>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   MC_WG_LDS_MAX) == hipSuccess &&
               hipFuncSetAttribute(reinterpret_cast<const void*>(&mc_wg_kernel<1024>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   MC_WG_LDS_MAX) == hipSuccess)
                ok = 1;
            (void)hipGetLastError();
        }
        return ok == 1;
    }

    // Greedy first-fit colouring over the symmetrised pattern: rows taken in
    // `order`, each the smallest colour no already coloured neighbour holds.
    // Neighbours of row i: ci[rp[i] .. rp[i+1]) and tci[trp[i] .. trp[i+1])
    // (the transpose, so that a_ji != 0 couples i and j as well).
    int colour_greedy(const std::vector<int>& rp, const std::vector<int>& ci,
                      const std::vector<int>& trp, const std::vector<int>& tci,
                      const std::vector<int>& order, std::vector<int>& colour)
    {
        const int m = (int)order.size();
        colour.assign(m, -1);
        std::vector<int> seen; // seen[c] == i: colour c is taken by a neighbour of i
        int              ncol = 0;
        for(int i : order)
        {
            auto mark = [&](int j) {
                if(j != i && colour[j] >= 0)
                    seen[colour[j]] = i;
            };
            for(int k = rp[i]; k < rp[i + 1]; k++)
                mark(ci[k]);
            for(int k = trp[i]; k < trp[i + 1]; k++)
                mark(tci[k]);
            int c = 0;
            while(c < ncol && seen[c] == i)
                c++;
            if(c == ncol)
            {
                ncol++;
                seen.push_back(-1);
            }
            colour[i] = c;
        }
        return ncol;
    }

    // Smallest-last order (Matula and Beck): repeatedly remove a row of least
    // remaining degree; colour in the reverse of the removal order.  Degrees
    // count the symmetrised pattern's entries (a coupling present in both A
    // and its transpose counts twice); buckets by degree as linked lists, so
    // the whole order is O(m + nnz).
    std::vector<int> smallest_last(const std::vector<int>& rp, const std::vector<int>& ci,
                                   const std::vector<int>& trp, const std::vector<int>& tci)
    {
        const int        m = (int)rp.size() - 1;
        std::vector<int> deg(m), next(m), prev(m);
        int              maxd = 0;
        for(int i = 0; i < m; i++)
        {
            int d = 0;
            for(int k = rp[i]; k < rp[i + 1]; k++)
                d += ci[k] != i;
            for(int k = trp[i]; k < trp[i + 1]; k++)
                d += tci[k] != i;
            deg[i] = d;
            maxd   = std::max(maxd, d);
        }
        std::vector<int> head(maxd + 1, -1);
        auto             unlink = [&](int v) {
            if(prev[v] >= 0)
                next[prev[v]] = next[v];
            else
                head[deg[v]] = next[v];
            if(next[v] >= 0)
                prev[next[v]] = prev[v];
        };
        auto push = [&](int v) {
            prev[v] = -1;
            next[v] = head[deg[v]];
            if(next[v] >= 0)
                prev[next[v]] = v;
            head[deg[v]] = v;
        };
        for(int i = m - 1; i >= 0; i--)
            push(i);
        std::vector<char> gone(m, 0);
        std::vector<int>  order(m);
        int               lo = 0;
        for(int k = m - 1; k >= 0; k--)
        {
            while(head[lo] < 0) // lo never exceeds the least remaining degree
                lo++;
            const int v = head[lo];
            unlink(v);
            gone[v]  = 1;
            order[k] = v;
            auto drop = [&](int u) {
                if(u != v && !gone[u])
                {
                    unlink(u);
                    deg[u]--;
                    push(u);
                    lo = std::min(lo, deg[u]);
                }
            };
            for(int e = rp[v]; e < rp[v + 1]; e++)
                drop(ci[e]);
            for(int e = trp[v]; e < trp[v + 1]; e++)
                drop(tci[e]);
        }
        return order;
    }

    // Colours of the rows: greedy in row order (the red-black colouring of a
    // 5-/7-point grid operator in lexicographic order), or greedy in
    // smallest-last order when that needs fewer colours -- on the coarse
    // Galerkin operators of a 128^3 Poisson hierarchy 54 / 72 / 60 against
    // 60 / 83 / 76 on levels 2-4, each colour one launch or one step of a
    // sweep.
    int colour_rows(const std::vector<int>& rp, const std::vector<int>& ci, int m,
                    std::vector<int>& colour)
    {
        const int nnz = rp[m];
        std::vector<int> trp(m + 1, 0), tci(nnz);
        for(int k = 0; k < nnz; k++)
            trp[ci[k] + 1]++;
        for(int i = 0; i < m; i++)
            trp[i + 1] += trp[i];
        std::vector<int> fill(trp.begin(), trp.end() - 1);
        for(int i = 0; i < m; i++)
            for(int k = rp[i]; k < rp[i + 1]; k++)
                tci[fill[ci[k]]++] = i;
        std::vector<int> natural(m);
        for(int i = 0; i < m; i++)
            natural[i] = i;
        int ncol = colour_greedy(rp, ci, trp, tci, natural, colour);
        if(ncol > 2)
        {
            std::vector<int> alt;
            const int        n2 = colour_greedy(rp, ci, trp, tci, smallest_last(rp, ci, trp, tci), alt);
            if(n2 < ncol)
            {
                colour.swap(alt);
                ncol = n2;
            }
        }
        return ncol;
    }

    int lanes_for(double mean_row)
    {
        int l = 1;
        while(l < 64 && 4 * l < mean_row)
            l <<= 1;
        return l;
    }

    // The one-workgroup sweep's plan (mc_wg_kernel) of a colour-ordered level
    // when it qualifies: at most max_nnz entries (one workgroup streams the
    // matrix from memory far slower than the per-colour launches on larger
    // levels), no row longer than a chunk, and x, b, the product buffers and
    // the plan within one workgroup's LDS.  NT is the smallest of 64, 256 and 1024 lanes
    // whose chunk holds the level's mean colour.  Every chunk is checked
    // against what the kernel assumes (a level whose plan failed that check
    // would take the per-colour launches).
    void mc_wg_plan(mc_sched& s, const std::vector<int>& prp, int max_nnz)
    {
        const int m  = (int)prp.size() - 1;
        s.wg         = false;
        s.nchunks    = 0;
        s.plan_h.clear();
        if(m <= 0 || s.nnz <= 0 || s.nnz > max_nnz || s.ncol <= 0)
            return;
        int longest = 0;
        for(int t = 0; t < m; t++)
            longest = std::max(longest, prp[t + 1] - prp[t]);
        const double mean_colour = (double)s.nnz / s.ncol;
        int          nt          = 64;
        while(nt < 1024 && (mean_colour > nt * MC_WG_U || longest > nt * MC_WG_U))
            nt *= 4;
        if(longest > nt * MC_WG_U)
            return;
        int lg = 0; // lanes per row: the level's, at most 16 (group_sum16)
        while((1 << lg) < s.lanes && lg < 4)
            lg++;
        const int ch = nt * MC_WG_U, rows_max = nt >> lg;
        std::vector<int>& plan = s.plan_h;
        for(int pass = 0; pass < 2; pass++)
        {
            const size_t first = plan.size();
            for(int q = 0; q < s.ncol; q++)
            {
                const int k  = pass == 0 ? q : s.ncol - 1 - q;
                const int t1 = s.off[k + 1];
                for(int t = s.off[k]; t < t1;)
                {
                    const int ta = t, c0 = prp[ta];
                    while(t < t1 && t - ta < rows_max && prp[t + 1] - c0 <= ch)
                        t++;
                    plan.insert(plan.end(), {c0, prp[t], ta, t, t == t1 ? 1 : 0, 0, 0, 0});
                }
            }
            while((plan.size() - first) / 8 % MC_WG_PF != 0)
                plan.insert(plan.end(), {s.nnz, s.nnz, 0, 0, 1, 0, 0, 0});
            if(pass == 0)
                s.nchunks = (int)(plan.size() / 8);
        }
        if((int)(plan.size() / 8) != 2 * s.nchunks ||
           mc_wg_lds_bytes(nt, m, 2 * s.nchunks) > MC_WG_LDS_MAX)
            return;
        for(size_t j = 0; j < plan.size(); j += 8)
        {
            const int c0 = plan[j], c1 = plan[j + 1], ta = plan[j + 2], tb = plan[j + 3];
            const bool ok = 0 <= c0 && c0 <= c1 && c1 <= s.nnz && c1 - c0 <= ch && 0 <= ta &&
                            ta <= tb && tb <= m && tb - ta <= rows_max &&
                            (ta == tb || (prp[ta] == c0 && prp[tb] == c1));
            if(!ok)
                return;
        }
        s.wg_nt = nt;
        s.wg_lg = lg;
        s.wg    = true;
    }
}

namespace lal
{
    int build_mc(lal_context* c, const csr_dev& A, mc_sched& s)
    {
        const int        m = A.m;
        std::vector<int> rp(m + 1), ci(std::max(A.nnz, 1));
        std::vector<double> v(std::max(A.nnz, 1));
        LAL_HIP(hipMemcpyAsync(rp.data(), A.rp, sizeof(int) * (m + 1), hipMemcpyDeviceToHost,
                               c->stream));
        if(A.nnz)
        {
            LAL_HIP(hipMemcpyAsync(ci.data(), A.ci, sizeof(int) * A.nnz, hipMemcpyDeviceToHost,
                                   c->stream));
            LAL_HIP(hipMemcpyAsync(v.data(), A.v, sizeof(double) * A.nnz,
                                   hipMemcpyDeviceToHost, c->stream));
        }
        LAL_HIP(hipStreamSynchronize(c->stream));
        std::vector<int> colour;
        const int        ncol = colour_rows(rp, ci, m, colour);
        s.off.assign(ncol + 1, 0);
        for(int i = 0; i < m; i++)
            s.off[colour[i] + 1]++;
        for(int k = 0; k < ncol; k++)
            s.off[k + 1] += s.off[k];
        std::vector<int> perm(m), fill(s.off.begin(), s.off.end() - 1);
        for(int i = 0; i < m; i++)
            perm[fill[colour[i]]++] = i;
        std::vector<int>    prp(m + 1, 0), pci;
        std::vector<double> pv, pdinv(m, 0.0);
        pci.reserve(A.nnz);
        pv.reserve(A.nnz);
        for(int t = 0; t < m; t++)
        {
            const int i = perm[t];
            double    d = 0.0;
            for(int k = rp[i]; k < rp[i + 1]; k++)
            {
                if(ci[k] == i)
                {
                    if(d == 0.0)
                        d = v[k]; // the first diagonal entry, as the level's 1/a_ii
                    continue;
                }
                pci.push_back(ci[k]);
                pv.push_back(v[k]);
            }
            if(d == 0.0)
            {
                set_last_error("multicolour Gauss-Seidel: a row without a nonzero diagonal");
                return LAL_ERR_UNSUPPORTED;
            }
            pdinv[t]   = 1.0 / d;
            prp[t + 1] = (int)pci.size();
        }
        s.ncol  = ncol;
        s.lanes = lanes_for(m ? (double)pci.size() / m : 0.0);
        s.nnz = (int)pci.size();
        mc_wg_plan(s, prp, MC_WG_MAX_NNZ);
        const std::vector<int>& plan = s.plan_h;
        if(s.wg)
            LAL_TRY(alloc_n(&s.plan, plan.size()));
        LAL_TRY(alloc_n(&s.perm, m));
        LAL_TRY(alloc_n(&s.off_d, s.off.size()));
        LAL_TRY(alloc_n(&s.prp, m + 1));
        LAL_TRY(alloc_n(&s.pci, pci.size()));
        LAL_TRY(alloc_n(&s.pv, pv.size()));
        LAL_TRY(alloc_n(&s.pdinv, m));
        LAL_HIP(hipMemcpyAsync(s.perm, perm.data(), sizeof(int) * m, hipMemcpyHostToDevice,
                               c->stream));
        LAL_HIP(hipMemcpyAsync(s.off_d, s.off.data(), sizeof(int) * s.off.size(),
                               hipMemcpyHostToDevice, c->stream));
        LAL_HIP(hipMemcpyAsync(s.prp, prp.data(), sizeof(int) * (m + 1), hipMemcpyHostToDevice,
                               c->stream));
        if(!pci.empty())
        {
            LAL_HIP(hipMemcpyAsync(s.pci, pci.data(), sizeof(int) * pci.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_HIP(hipMemcpyAsync(s.pv, pv.data(), sizeof(double) * pv.size(),
                                   hipMemcpyHostToDevice, c->stream));
        }
        LAL_HIP(hipMemcpyAsync(s.pdinv, pdinv.data(), sizeof(double) * m, hipMemcpyHostToDevice,
                               c->stream));
        if(s.wg)
            LAL_HIP(hipMemcpyAsync(s.plan, plan.data(), sizeof(int) * plan.size(),
                                   hipMemcpyHostToDevice, c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        s.built = true;
        return LAL_OK;
    }

    // One multicolour sweep on x in place: forward colour order, then
    // backward when sym.
    int mc_sweep(lal_context* c, const mc_sched& s, int m, const double* b, double* x, bool sym)
    {
        if(m == 0 || s.ncol == 0)
            return LAL_OK;
        if(s.wg && mc_wg_lds_ok())
        {
            // forward: the plan's first half; symmetric: both halves
            const int steps = s.nchunks * (sym ? 2 : 1);
            const int lds   = (int)mc_wg_lds_bytes(s.wg_nt, m, steps);
            const auto* pl  = reinterpret_cast<const int4v*>(s.plan);
#define LAL_MC_WG(NT)                                                                      \
    mc_wg_kernel<NT><<<1, NT, lds, c->stream>>>(m, s.nnz, steps, s.wg_lg, pl, s.perm, s.prp, \
                                                s.pci, s.pv, s.pdinv, b, x)
            switch(s.wg_nt)
            {
            case 64: LAL_MC_WG(64); break;
            case 256: LAL_MC_WG(256); break;
            default: LAL_MC_WG(1024); break;
            }
#undef LAL_MC_WG
            LAL_LAUNCHED();
            return LAL_OK;
        }
        for(int pass = 0; pass < (sym ? 2 : 1); pass++)
            for(int q = 0; q < s.ncol; q++)
            {
                const int k   = pass == 0 ? q : s.ncol - 1 - q;
                const int t0  = s.off[k];
                const int cnt = s.off[k + 1] - t0;
                if(cnt == 0)
                    continue;
                const long long threads = (long long)cnt * s.lanes;
                const int       grid    = (int)((threads + 255) / 256);
#define LAL_MC_LAUNCH(L)                                                                   \
    mc_colour_kernel<L><<<grid, 256, 0, c->stream>>>(cnt, s.perm + t0, s.prp + t0, s.pci,    \
                                                     s.pv, s.pdinv + t0, b, x)
                switch(s.lanes)
                {
                case 1: LAL_MC_LAUNCH(1); break;
                case 2: LAL_MC_LAUNCH(2); break;
                case 4: LAL_MC_LAUNCH(4); break;
                case 8: LAL_MC_LAUNCH(8); break;
                case 16: LAL_MC_LAUNCH(16); break;
                case 32: LAL_MC_LAUNCH(32); break;
                default: LAL_MC_LAUNCH(64); break;
                }
#undef LAL_MC_LAUNCH
                LAL_LAUNCHED();
            }
        return LAL_OK;
    }
}
