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

    // A whole sweep in one workgroup: the level sets in order, a barrier
    // between them (the rows written in one set are read by later sets of the
    // same workgroup), for operators whose sets are small.
    __global__ void __launch_bounds__(1024)
        gs_sweep_wg_kernel(int nsets, const int* __restrict__ off, const int* __restrict__ perm,
                           const int* __restrict__ rp, const int* __restrict__ ci,
                           const double* __restrict__ v, const double* __restrict__ b, double* x,
                           int sor)
    {
        for(int s = 0; s < nsets; s++)
        {
            const int e = off[s + 1];
            for(int t = off[s] + (int)threadIdx.x; t < e; t += 1024)
                gs_row(perm[t], rp, ci, v, b, x, sor);
            __syncthreads();
        }
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

    int build_sched(lal_context* c, const csr_dev& A, gs_sched& f, gs_sched& bw)
    {
        std::vector<int> rp(A.m + 1), ci(std::max(A.nnz, 1));
        LAL_HIP(hipMemcpyAsync(rp.data(), A.rp, sizeof(int) * (A.m + 1), hipMemcpyDeviceToHost,
                               c->stream));
        if(A.nnz)
            LAL_HIP(hipMemcpyAsync(ci.data(), A.ci, sizeof(int) * A.nnz, hipMemcpyDeviceToHost,
                                   c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        for(int dir = 0; dir < 2; dir++)
        {
            gs_sched&        g = dir == 0 ? f : bw;
            std::vector<int> perm;
            level_sets(rp, ci, A.m, dir == 0, perm, g.off);
            LAL_TRY(lal::alloc_n(&g.perm, perm.size()));
            LAL_HIP(hipMemcpyAsync(g.perm, perm.data(), sizeof(int) * perm.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_TRY(lal::alloc_n(&g.off_d, g.off.size()));
            LAL_HIP(hipMemcpyAsync(g.off_d, g.off.data(), sizeof(int) * g.off.size(),
                                   hipMemcpyHostToDevice, c->stream));
        }
        LAL_HIP(hipStreamSynchronize(c->stream));
        return LAL_OK;
    }

    int gs_sweep(lal_context* c, const csr_dev& A, const gs_sched& g, const double* b, double* x,
                 int sor)
    {
        const int nsets = (int)g.off.size() - 1;
        if(nsets > 0 && A.m / nsets < 1024)
        {
            gs_sweep_wg_kernel<<<1, 1024, 0, c->stream>>>(nsets, g.off_d, g.perm, A.rp, A.ci, A.v,
                                                          b, x, sor);
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
        if(!L.gs_fwd.perm)
            LAL_TRY(build_sched(c, A, L.gs_fwd, L.gs_bwd));
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
        LAL_HIP(hipMemsetAsync(Wc.x, 0, sizeof(double) * R.m, c->stream));
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

extern "C" int lal_amg_solve(lal_handle c, lal_hierarchy h, const double* b, double* x, int n1,
                             int n2, int cycle, int smoother, const lal_solve_options* opts,
                             lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    LAL_REQUIRE(n1 >= 0 && n2 >= 0 && cycle >= 0 && cycle <= 2 && smoother >= 0 &&
                smoother <= 4);
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
        LAL_TRY(cycle_on(h, 0, x, b, n1, n2, n2, cycle, smoother));
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
