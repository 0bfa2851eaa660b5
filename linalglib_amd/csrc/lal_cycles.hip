// AMG as a solver on the device: V, W and F cycles with the Jacobi,
// Gauss-Seidel, "symmetric" Gauss-Seidel, SOR and SSOR smoothers.
//
//   apply_smoother      iterative_solvers/amg/amg.cpp:110-142 (SOR/SSOR with
//                       omega = 0.5; the symmetric Gauss-Seidel case runs a
//                       forward Gauss-Seidel sweep, as written there)
//   sweeps              iterative_solvers/classic/gauss_seidel.cpp:50-86,
//                       sor.cpp:43-83, ssor.cpp:43-111
//   vcycle/wcycle/fcycle amg.cpp:144-513 (coarsest: the 1x1 divide or 1000
//                       Gauss-Seidel sweeps for V, 100 sweeps for W and F)
//   amg_solve           amg.cpp:516-571 with iter_control.cpp:38-91
//
// Gauss-Seidel is sequential by definition; on the device each sweep runs as
// level sets of the rows' dependency graph (row j must follow every adjacent
// row i < j, whose new value it reads, and precede every adjacent row k > j,
// whose old value it reads).  Rows of one set are independent, one launch per
// set, and every row is updated with the reference's own arithmetic, so a
// sweep is bit-identical to the sequential one.  With a hierarchy built in
// exact mode the whole solve is bit-identical to the reference's; in the fast
// mode the residuals, transfers and Jacobi sweeps reorder sums inside rows.
#include "lal_hier.h"

#include <cstring>
#include <vector>

namespace
{
    using lal::csr_dev;
    using lal::gs_sched;

    // One Gauss-Seidel (sor = 0) or SOR (omega = 1/2) update of row j: sigma
    // sums the off-diagonal products with a separate multiply and add in CSR
    // order, a_jj is the row's diagonal entry.
    __device__ __forceinline__ void gs_row(int j, const int* __restrict__ rp,
                                           const int* __restrict__ ci,
                                           const double* __restrict__ v,
                                           const double* __restrict__ b, double* x, int sor)
    {
        const int rs    = rp[j];
        const int re    = rp[j + 1];
        double    sigma = 0.0, ajj = 0.0;
        // eight entries' loads in flight at a time; the sum stays in CSR order
        for(int k0 = rs; k0 < re; k0 += 8)
        {
            int    col[8];
            double vv[8], xx[8];
#pragma unroll
            for(int u = 0; u < 8; u++)
            {
                const int k = min(k0 + u, re - 1);
                col[u]      = ci[k];
                vv[u]       = v[k];
            }
#pragma unroll
            for(int u = 0; u < 8; u++)
                xx[u] = x[col[u]];
#pragma unroll
            for(int u = 0; u < 8; u++)
            {
                if(k0 + u >= re)
                    break;
                if(col[u] != j)
                    sigma = sigma + vv[u] * xx[u];
                else
                    ajj = vv[u];
            }
        }
        x[j] = sor ? x[j] + 0.5 * ((b[j] - sigma) / ajj - x[j]) : (b[j] - sigma) / ajj;
    }

    // The rows perm[0 .. cnt) of one level set, one lane each.
    __global__ void __launch_bounds__(256)
        gs_level_kernel(int cnt, const int* __restrict__ perm, const int* __restrict__ rp,
                        const int* __restrict__ ci, const double* __restrict__ v,
                        const double* __restrict__ b, double* x, int sor)
    {
        const int t = blockIdx.x * 256 + threadIdx.x;
        if(t < cnt)
            gs_row(perm[t], rp, ci, v, b, x, sor);
    }

    // A whole sweep in one workgroup for operators whose level sets are thin
    // (few rows, long rows).  The sweep is a host-built list of tiles: a
    // tile is up to GS_CH consecutive entries of a group of at most 1024 rows
    // of one set, in the copy of the matrix whose rows are in set order.
    // Per tile the lanes form the products a_jk x_k of the tile's entries
    // into LDS (+0.0 in a diagonal entry's slot); then one lane per row adds
    // its row's slots in CSR order.  sigma starts at +0.0 and so is never
    // -0.0, which makes adding the diagonal's +0.0 an exact no-op: the sums
    // are those of the sequential sweep, which skips the diagonal.  a_jj (the
    // row's last diagonal entry, 0 if none) comes from pdiag.  The rows are
    // written after their group's last tile and a barrier ends each set.
    // What does not depend on x (the next tile's columns and values, the next
    // group's row bounds, b) is loaded before the LDS-only barriers, so a
    // tile waits only for its gathers of x.  With XLDS, x lives in LDS.
    constexpr int GS_CH       = 8192;
    constexpr int GS_PER_LANE = GS_CH / 1024;
    // tile flags (low 3 bits of the fourth word; the group's end row above)
    constexpr int GS_START = 1, GS_END_GROUP = 2, GS_END_SET = 4;

    using lal::lds_barrier;

    __device__ __forceinline__ void gs_tile_loads(int c0, int c1, int tid,
                                                  const int* __restrict__ pci,
                                                  const double* __restrict__ pv, int* col,
                                                  double* vv)
    {
#pragma unroll
        for(int u = 0; u < GS_PER_LANE; u++)
        {
            // clamped into the tile (entry 0 for an empty one) instead of a branch
            const int k = max(0, min(c0 + tid + u * 1024, c1 - 1));
            col[u]      = pci[k];
            vv[u]       = pv[k];
        }
    }

    template <bool XLDS>
    __global__ void __launch_bounds__(1024)
        gs_sweep_wg_kernel(int m, int ntiles, const int4* __restrict__ tiles,
                           const int* __restrict__ perm, const int* __restrict__ prp,
                           const int* __restrict__ pci, const double* __restrict__ pv,
                           const double* __restrict__ pdiag, const double* __restrict__ b,
                           double* x, int sor)
    {
        __shared__ double        s_p[GS_CH];
        extern __shared__ double s_x[];
        double*                  X   = XLDS ? s_x : x;
        const int                tid = threadIdx.x;
        if(XLDS)
        {
            for(int i = tid; i < m; i += 1024)
                s_x[i] = x[i];
            __syncthreads();
        }
        int4   T = tiles[0];
        int    col[GS_PER_LANE];
        double vv[GS_PER_LANE];
        gs_tile_loads(T.x, T.y, tid, pci, pv, col, vv);
        int    j     = T.z + tid;
        bool   mine  = j < (T.w >> 3);
        int    rs    = mine ? prp[j] : 0;
        int    re    = mine ? prp[j + 1] : 0;
        int    row   = mine ? perm[j] : 0;
        double ajj   = mine ? pdiag[j] : 1.0;
        double sigma = 0.0;
        for(int t = 0; t < ntiles; t++)
        {
            const int  fl   = T.w & 7;
            const bool more = t + 1 < ntiles;
            const int4 N    = more ? tiles[t + 1] : T;
            double     xv[GS_PER_LANE];
#pragma unroll
            for(int u = 0; u < GS_PER_LANE; u++)
                xv[u] = X[col[u] < 0 ? 0 : col[u]];
            double bj = 0.0, xo = 0.0;
            if((fl & GS_END_GROUP) && mine)
            {
                bj = b[row];
                xo = X[row]; // written by this row only
            }
            // the next tile's x-independent loads, in flight across the barriers
            int    ncol[GS_PER_LANE];
            double nvv[GS_PER_LANE];
            gs_tile_loads(N.x, N.y, tid, pci, pv, ncol, nvv);
            const bool ngrp  = more && (N.w & GS_START);
            const int  nj    = N.z + tid;
            const bool nmine = ngrp ? nj < (N.w >> 3) : mine;
            int        nrs = rs, nre = re, nrow = row;
            double     najj = ajj;
            if(ngrp && nmine)
            {
                nrs  = prp[nj];
                nre  = prp[nj + 1];
                nrow = perm[nj];
                najj = pdiag[nj];
            }
#pragma unroll
            for(int u = 0; u < GS_PER_LANE; u++)
            {
                const int k = T.x + tid + u * 1024;
                if(k < T.y)
                    s_p[k - T.x] = col[u] < 0 ? 0.0 : vv[u] * xv[u];
            }
            lds_barrier();
            if(mine)
            {
                const int k0 = max(rs, T.x) - T.x;
                const int k1 = min(re, T.y) - T.x;
                int       k  = k0;
                for(; k + 8 <= k1; k += 8)
                {
                    double p[8];
#pragma unroll
                    for(int u = 0; u < 8; u++)
                        p[u] = s_p[k + u];
#pragma unroll
                    for(int u = 0; u < 8; u++)
                        sigma = sigma + p[u];
                }
                for(; k < k1; k++)
                    sigma = sigma + s_p[k];
            }
            lds_barrier();
            if(fl & GS_END_GROUP)
            {
                if(mine)
                    X[row] = sor ? xo + 0.5 * ((bj - sigma) / ajj - xo) : (bj - sigma) / ajj;
                sigma = 0.0;
            }
            if(fl & GS_END_SET)
            {
                if(XLDS)
                    lds_barrier();
                else
                    __syncthreads();
            }
            T = N;
#pragma unroll
            for(int u = 0; u < GS_PER_LANE; u++)
            {
                col[u] = ncol[u];
                vv[u]  = nvv[u];
            }
            if(ngrp)
            {
                mine = nmine;
                rs   = nrs;
                re   = nre;
                row  = nrow;
                ajj  = najj;
            }
        }
        if(XLDS)
        {
            __syncthreads();
            for(int i = tid; i < m; i += 1024)
                x[i] = s_x[i];
        }
    }

    // The tiles of one sweep from the set offsets and the set-ordered row
    // pointer; false when a group end row does not fit beside the flags.
    bool gs_tiles(const std::vector<int>& off, const std::vector<int>& prp,
                  std::vector<int>& tiles)
    {
        tiles.clear();
        for(size_t s = 0; s + 1 < off.size(); s++)
        {
            const int a = off[s], e = off[s + 1];
            for(int g0 = a; g0 < e; g0 += 1024)
            {
                const int g1 = std::min(e, g0 + 1024);
                if(g1 >= (1 << 28))
                    return false;
                const int e0     = prp[g0];
                const int e1     = prp[g1];
                const int chunks = std::max(1, (e1 - e0 + GS_CH - 1) / GS_CH);
                for(int ch = 0; ch < chunks; ch++)
                {
                    const int c0 = e0 + ch * GS_CH;
                    const int c1 = std::min(e1, c0 + GS_CH);
                    const int fl = (ch == 0 ? GS_START : 0) |
                                   (ch == chunks - 1 ? GS_END_GROUP : 0) |
                                   (ch == chunks - 1 && g1 == e ? GS_END_SET : 0);
                    tiles.insert(tiles.end(), {c0, c1, g0, (g1 << 3) | fl});
                }
            }
        }
        return !tiles.empty();
    }

    // Bytes of LDS one workgroup may hold on the current device, and the
    // dynamic part the single-workgroup sweep may take for x.
    int wg_sweep_lds_limit(int* dyn_max)
    {
        static int cached = -1;
        if(cached < 0)
        {
            int dev = 0, v = 0;
            LAL_HIP(hipGetDevice(&dev));
            LAL_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
            const int stat = GS_CH * (int)sizeof(double);
            int       dyn  = v - stat;
            if(dyn > 0 && v > 65536 &&
               hipFuncSetAttribute(reinterpret_cast<const void*>(&gs_sweep_wg_kernel<true>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   dyn) != hipSuccess)
            {
                (void)hipGetLastError();
                dyn = 65536 - stat;
            }
            cached = dyn > 0 ? dyn : 0;
        }
        *dyn_max = cached;
        return LAL_OK;
    }

    // Level sets of the forward (and backward) sweep from the host pattern:
    // a row's set is one more than the largest set of the adjacent rows it
    // must follow.  Pattern entries in either triangle count, so patterns
    // that are not symmetric are ordered correctly too.
    void level_sets(const std::vector<int>& rp, const std::vector<int>& ci, int m, bool forward,
                    std::vector<int>& perm, std::vector<int>& off)
    {
        std::vector<int> lev(m, 0);
        int              top = 0;
        for(int s = 0; s < m; s++)
        {
            const int j = forward ? s : m - 1 - s;
            int       l = lev[j];
            for(int k = rp[j]; k < rp[j + 1]; k++)
            {
                const int c = ci[k];
                if(forward ? c < j : c > j)
                    l = std::max(l, lev[c] + 1);
            }
            lev[j] = l;
            for(int k = rp[j]; k < rp[j + 1]; k++)
            {
                const int c = ci[k];
                if(forward ? c > j : c < j)
                    lev[c] = std::max(lev[c], l + 1);
            }
            top = std::max(top, l);
        }
        off.assign(top + 2, 0);
        for(int j = 0; j < m; j++)
            off[lev[j] + 1]++;
        for(int k = 0; k <= top; k++)
            off[k + 1] += off[k];
        perm.assign(m, 0);
        std::vector<int> fill(off.begin(), off.end() - 1);
        for(int j = 0; j < m; j++)
            perm[fill[lev[j]]++] = j;
    }

    // Whether a sweep over these sets runs in one workgroup.
    bool wg_sweep(int m, size_t nsets)
    {
        return nsets > 0 && (size_t)m / nsets < 1024;
    }

    // Level sets of one sweep direction; for a single-workgroup sweep also the
    // set-ordered copy of the matrix.
    int build_sched(lal_context* c, const csr_dev& A, bool forward, gs_sched& g)
    {
        std::vector<int> rp(A.m + 1), ci(std::max(A.nnz, 1));
        LAL_HIP(hipMemcpyAsync(rp.data(), A.rp, sizeof(int) * (A.m + 1), hipMemcpyDeviceToHost,
                               c->stream));
        if(A.nnz)
            LAL_HIP(hipMemcpyAsync(ci.data(), A.ci, sizeof(int) * A.nnz, hipMemcpyDeviceToHost,
                                   c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        std::vector<int> perm;
        level_sets(rp, ci, A.m, forward, perm, g.off);
        LAL_TRY(lal::alloc_n(&g.perm, perm.size()));
        LAL_HIP(hipMemcpyAsync(g.perm, perm.data(), sizeof(int) * perm.size(),
                               hipMemcpyHostToDevice, c->stream));
        LAL_TRY(lal::alloc_n(&g.off_d, g.off.size()));
        LAL_HIP(hipMemcpyAsync(g.off_d, g.off.data(), sizeof(int) * g.off.size(),
                               hipMemcpyHostToDevice, c->stream));
        if(wg_sweep(A.m, g.off.size() - 1) && A.nnz > 0)
        {
            std::vector<double> v(A.nnz);
            LAL_HIP(hipMemcpyAsync(v.data(), A.v, sizeof(double) * A.nnz,
                                   hipMemcpyDeviceToHost, c->stream));
            LAL_HIP(hipStreamSynchronize(c->stream));
            std::vector<int>    prp(A.m + 1, 0), pci(A.nnz);
            std::vector<double> pv(A.nnz), pdiag(A.m, 0.0);
            for(int t = 0; t < A.m; t++)
            {
                const int j = perm[t];
                int       o = prp[t];
                for(int k = rp[j]; k < rp[j + 1]; k++, o++)
                {
                    pci[o] = ci[k] == j ? -1 - ci[k] : ci[k];
                    pv[o]  = v[k];
                    if(ci[k] == j)
                        pdiag[t] = v[k]; // the last one, as the sweep's loop keeps
                }
                prp[t + 1] = o;
            }
            std::vector<int> tiles;
            if(!gs_tiles(g.off, prp, tiles))
            {
                LAL_HIP(hipStreamSynchronize(c->stream));
                return LAL_OK; // swept one launch per set instead
            }
            g.ntiles = (int)(tiles.size() / 4);
            LAL_TRY(lal::alloc_n(&g.tiles, tiles.size()));
            LAL_HIP(hipMemcpyAsync(g.tiles, tiles.data(), sizeof(int) * tiles.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_TRY(lal::alloc_n(&g.prp, prp.size()));
            LAL_TRY(lal::alloc_n(&g.pci, pci.size()));
            LAL_TRY(lal::alloc_n(&g.pv, pv.size()));
            LAL_HIP(hipMemcpyAsync(g.prp, prp.data(), sizeof(int) * prp.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_HIP(hipMemcpyAsync(g.pci, pci.data(), sizeof(int) * pci.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_HIP(hipMemcpyAsync(g.pv, pv.data(), sizeof(double) * pv.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_TRY(lal::alloc_n(&g.pdiag, pdiag.size()));
            LAL_HIP(hipMemcpyAsync(g.pdiag, pdiag.data(), sizeof(double) * pdiag.size(),
                                   hipMemcpyHostToDevice, c->stream));
        }
        LAL_HIP(hipStreamSynchronize(c->stream));
        return LAL_OK;
    }

    int gs_sweep(lal_context* c, const csr_dev& A, const gs_sched& g, const double* b, double* x,
                 int sor)
    {
        if(g.prp)
        {
            int dyn_max = 0;
            LAL_TRY(wg_sweep_lds_limit(&dyn_max));
            const size_t xb    = sizeof(double) * (size_t)A.m;
            const int4*  tiles = reinterpret_cast<const int4*>(g.tiles);
            if(xb <= (size_t)dyn_max)
                gs_sweep_wg_kernel<true><<<1, 1024, xb, c->stream>>>(
                    A.m, g.ntiles, tiles, g.perm, g.prp, g.pci, g.pv, g.pdiag, b, x, sor);
            else
                gs_sweep_wg_kernel<false><<<1, 1024, 0, c->stream>>>(
                    A.m, g.ntiles, tiles, g.perm, g.prp, g.pci, g.pv, g.pdiag, b, x, sor);
            LAL_LAUNCHED();
            return LAL_OK;
        }
        for(size_t k = 0; k + 1 < g.off.size(); k++)
        {
            const int cnt = g.off[k + 1] - g.off[k];
            if(cnt == 0)
                continue;
            gs_level_kernel<<<(cnt + 255) / 256, 256, 0, c->stream>>>(cnt, g.perm + g.off[k], A.rp,
                                                                      A.ci, A.v, b, x, sor);
            LAL_LAUNCHED();
        }
        return LAL_OK;
    }

    // apply_smoother on level l, x in place.
    int smooth(lal_hierarchy_t* h, int l, double* x, const double* b, int smoother)
    {
        lal_context*        c = h->c;
        lal::level_t&       L = h->L[l];
        const csr_dev&      A = L.A;
        lal::work<double>&  W = L.w64;
        if(smoother == 0)
        {
            LAL_HIP(hipMemcpyAsync(W.t, x, sizeof(double) * A.m, hipMemcpyDeviceToDevice,
                                   c->stream));
            return lal::jacobi_sweep<double>(c, A.d, A.m, A.rp, A.ci, A.v, W.diag, W.dinv, W.t, b,
                                             x);
        }
        if(smoother >= 5)
        {
            if(!L.mc.built)
                LAL_TRY(lal::build_mc(c, A, L.mc));
            return lal::mc_sweep(c, L.mc, A.m, b, x, smoother == 6);
        }
        if(!L.gs_fwd.perm)
            LAL_TRY(build_sched(c, A, true, L.gs_fwd));
        if(smoother == 4 && !L.gs_bwd.perm)
            LAL_TRY(build_sched(c, A, false, L.gs_bwd));
        const int sor = smoother >= 3;
        LAL_TRY(gs_sweep(c, A, L.gs_fwd, b, x, sor));
        if(smoother == 4)
            LAL_TRY(gs_sweep(c, A, L.gs_bwd, b, x, sor));
        return LAL_OK;
    }

    // vcycle (cycle 0), wcycle (1), fcycle (2) on level l with x in place.
    int cycle_on(lal_hierarchy_t* h, int l, double* x, const double* b, int n1, int n2, int n3,
                 int cycle, int smoother)
    {
        lal_context*   c = h->c;
        const csr_dev& A = h->L[l].A;
        if(l == h->levels)
        {
            if(cycle == 0 && A.m == 1)
                return lal::coarse_scalar<double>(c, A.v, b, x);
            return lal::coarse_gs_sweeps(c, A.m, A.rp, A.ci, A.v, b, x,
                                         cycle == 0 ? 1000 : 100);
        }
        const csr_dev&     R  = h->L[l].R;
        const csr_dev&     P  = h->L[l].P;
        lal::work<double>& W  = h->L[l].w64;
        lal::work<double>& Wc = h->L[l + 1].w64;
        for(int i = 0; i < n1; i++)
            LAL_TRY(smooth(h, l, x, b, smoother));
        LAL_TRY(lal::residual_f64(c, A.d, A.m, A.n, A.nnz, A.rp, A.ci, A.v, x, b, W.r));
        LAL_TRY(lal::csrmv(c, R.d, R.m, R.n, R.nnz, 1.0, R.rp, R.ci, R.v, W.r, 0.0, Wc.b));
        LAL_TRY(lal_fill_f64(c, Wc.x, (size_t)R.m, 0.0)); // a kernel: the cycle may be captured
        LAL_TRY(cycle_on(h, l + 1, Wc.x, Wc.b, n1, n2, n3, cycle, smoother));
        LAL_TRY(lal::csrmv(c, P.d, P.m, P.n, P.nnz, 1.0, P.rp, P.ci, P.v, Wc.x, 1.0, x));
        for(int i = 0; i < n2; i++)
            LAL_TRY(smooth(h, l, x, b, smoother));
        if(cycle == 0)
            return LAL_OK;
        // second coarse visit: continues from the first one's correction,
        // which is added to x in full again (amg.cpp:322-341, 444-463)
        LAL_TRY(lal::residual_f64(c, A.d, A.m, A.n, A.nnz, A.rp, A.ci, A.v, x, b, W.r));
        LAL_TRY(lal::csrmv(c, R.d, R.m, R.n, R.nnz, 1.0, R.rp, R.ci, R.v, W.r, 0.0, Wc.b));
        LAL_TRY(cycle_on(h, l + 1, Wc.x, Wc.b, n1, n2, n3, cycle == 2 ? 0 : 1, smoother));
        LAL_TRY(lal::csrmv(c, P.d, P.m, P.n, P.nnz, 1.0, P.rp, P.ci, P.v, Wc.x, 1.0, x));
        for(int i = 0; i < n3; i++)
            LAL_TRY(smooth(h, l, x, b, smoother));
        return LAL_OK;
    }

    int residual_norm(lal_hierarchy_t* h, const double* x, const double* b, int slot, double* v)
    {
        lal_context*   c      = h->c;
        const csr_dev& A      = h->L[0].A;
        double*        part_r = h->part + 3 * lal::PCG_BLOCKS;
        unsigned       seq    = 0;
        LAL_TRY(lal::residual_f64(c, A.d, A.m, A.n, A.nnz, A.rp, A.ci, A.v, x, b, h->res));
        LAL_TRY(lal::norm_partials<double>(c, A.m, h->res, part_r));
        LAL_TRY(lal::post_scalar(h, 1, part_r, slot, &seq));
        return lal::wait_scalar(h, seq, slot, v);
    }
}

namespace
{
    // One cycle on level 0, x in place.  The multicolour smoothers (5, 6)
    // launch one kernel per colour and level, several hundred per cycle at
    // 128^3, each short enough that the host's launch rate bounds the cycle:
    // their cycles are captured once into a HIP graph and replayed.  A
    // (x, b, sweeps, cycle, smoother) combination runs eagerly the first
    // time -- which also builds every level's colouring, whose host copies
    // may not happen under capture -- and is captured on its second use and
    // replayed from then on.  A capture that fails leaves the combination
    // eager.  The graph holds kernels only (the cycle zeroes with a fill
    // kernel, not hipMemsetAsync) and is replayed on the caller's stream.
    int run_cycle(lal_hierarchy_t* h, double* x, const double* b, int n1, int n2, int n3,
                  int cycle, int smoother)
    {
        if(smoother < 5)
            return cycle_on(h, 0, x, b, n1, n2, n3, cycle, smoother);
        return lal::captured(h, x, b, n1, n2, n3, cycle, smoother,
                             [&] { return cycle_on(h, 0, x, b, n1, n2, n3, cycle, smoother); });
    }
}

namespace lal
{
    int vcycle_smoothed(lal_hierarchy_t* h, const double* r, double* z, int n1, int n2,
                        int smoother)
    {
        LAL_TRY(lal_fill_f64(h->c, z, (size_t)h->L[0].A.m, 0.0));
        return run_cycle(h, z, r, n1, n2, n2, 0, smoother);
    }
}

extern "C" int lal_amg_solve(lal_handle c, lal_hierarchy h, const double* b, double* x, int n1,
                             int n2, int cycle, int smoother, const lal_solve_options* opts,
                             lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    LAL_REQUIRE(n1 >= 0 && n2 >= 0 && cycle >= 0 && cycle <= 2 && smoother >= 0 &&
                smoother <= 6);
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.max_iter >= 0);
    double init = 0.0, rn = 0.0;
    LAL_TRY(residual_norm(h, x, b, lal::S_RNORM0, &init));
    rn         = init;
    int  count = 0;
    bool conv  = false;
    while(!(count > o.max_iter))
    {
        LAL_TRY(run_cycle(h, x, b, n1, n2, n2, cycle, smoother));
        LAL_TRY(residual_norm(h, x, b, lal::S_RNORM, &rn));
        if(lal::converges(rn, init, o.rel_tol, o.abs_tol))
        {
            conv = true;
            break;
        }
        count++;
    }
    if(info)
    {
        info->iterations       = count;
        info->outer_iterations = 0;
        info->converged        = conv;
        info->initial_res_norm = init;
        info->final_res_norm   = rn;
    }
    return LAL_OK;
}
