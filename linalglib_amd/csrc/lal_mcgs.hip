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
// The colouring is greedy first fit in row order (on the host, on first use of
// a level): row i takes the smallest colour that no already coloured row
// coupled to it holds.  On a 7-point grid operator in lexicographic order this
// is the red-black colouring.  The level keeps a copy of its matrix with the
// rows grouped by colour (off-diagonal entries only, and 1/a_ii per row), so
// a colour's rows are one contiguous stream.  Colours of many rows are one
// launch each, with 1 to 64 lanes per row by the level's mean row length;
// levels of at most MC_WG_ROWS rows sweep every colour in one workgroup with
// x in LDS and a barrier between colours.
#include "lal_hier.h"

#include <vector>

namespace
{
    using lal::csr_dev;
    using lal::mc_sched;

    constexpr int MC_WG_ROWS = 8192; // x of a one-workgroup level in LDS (64 KiB)

    // Rows t in [0, cnt) of one colour, LANES lanes per row.
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
        double    s  = 0.0;
        int       k  = rs + lig;
        for(; k + 3 * LANES < re; k += 4 * LANES)
        {
            int    c[4];
            double v[4], xx[4];
#pragma unroll
            for(int u = 0; u < 4; u++)
            {
                c[u] = pci[k + u * LANES];
                v[u] = pv[k + u * LANES];
            }
#pragma unroll
            for(int u = 0; u < 4; u++)
                xx[u] = x[c[u]];
#pragma unroll
            for(int u = 0; u < 4; u++)
                s += v[u] * xx[u];
        }
        for(; k < re; k += LANES)
            s += pv[k] * x[pci[k]];
#pragma unroll
        for(int o = LANES >> 1; o > 0; o >>= 1)
            s += __shfl_xor(s, o, 64);
        if(lig == 0)
        {
            const int i = perm[t];
            x[i]        = (b[i] - s) * pdinv[t];
        }
    }

    // Every colour of a small level in one workgroup of 1024 lanes, x in LDS
    // (m <= MC_WG_ROWS); forward colour order, then backward when `sym`.
    __global__ void __launch_bounds__(1024)
        mc_wg_kernel(int m, int ncol, const int* __restrict__ off, const int* __restrict__ perm,
                     const int* __restrict__ prp, const int* __restrict__ pci,
                     const double* __restrict__ pv, const double* __restrict__ pdinv,
                     const double* __restrict__ b, double* x, int lanes, int sym)
    {
        extern __shared__ double s_x[];
        const int tid = threadIdx.x;
        for(int i = tid; i < m; i += 1024)
            s_x[i] = x[i];
        __syncthreads();
        const int groups = 1024 / lanes;
        const int grp    = tid / lanes;
        const int lig    = tid % lanes;
        for(int pass = 0; pass < 1 + sym; pass++)
            for(int q = 0; q < ncol; q++)
            {
                const int k  = pass == 0 ? q : ncol - 1 - q;
                const int t0 = off[k], t1 = off[k + 1];
                for(int tb = t0; tb < t1; tb += groups)
                {
                    const int t = tb + grp;
                    double    s = 0.0;
                    if(t < t1)
                    {
                        const int re = prp[t + 1];
                        int       e  = prp[t] + lig;
                        for(; e + 3 * lanes < re; e += 4 * lanes)
                        {
                            int    c[4];
                            double v[4];
#pragma unroll
                            for(int u = 0; u < 4; u++)
                            {
                                c[u] = pci[e + u * lanes];
                                v[u] = pv[e + u * lanes];
                            }
#pragma unroll
                            for(int u = 0; u < 4; u++)
                                s += v[u] * s_x[c[u]];
                        }
                        for(; e < re; e += lanes)
                            s += pv[e] * s_x[pci[e]];
                    }
                    for(int o = lanes >> 1; o > 0; o >>= 1)
                        s += __shfl_xor(s, o, 64);
                    // rows of one colour are uncoupled: a row's sum reads no
                    // x that another row of its colour writes, so the writes
                    // of one colour need no barrier against its reads
                    if(t < t1 && lig == 0)
                    {
                        const int i = perm[t];
                        s_x[i]      = (b[i] - s) * pdinv[t];
                    }
                }
                __syncthreads();
            }
        for(int i = tid; i < m; i += 1024)
            x[i] = s_x[i];
    }

    // Greedy first-fit colouring in row order over the symmetrised pattern.
    int colour_rows(const std::vector<int>& rp, const std::vector<int>& ci, int m,
                    std::vector<int>& colour)
    {
        const int nnz = rp[m];
        // transpose pattern, so that a_ji != 0 couples i and j as well
        std::vector<int> trp(m + 1, 0), tci(nnz);
        for(int k = 0; k < nnz; k++)
            trp[ci[k] + 1]++;
        for(int i = 0; i < m; i++)
            trp[i + 1] += trp[i];
        std::vector<int> fill(trp.begin(), trp.end() - 1);
        for(int i = 0; i < m; i++)
            for(int k = rp[i]; k < rp[i + 1]; k++)
                tci[fill[ci[k]]++] = i;
        colour.assign(m, -1);
        std::vector<int> seen; // seen[c] == i: colour c is taken by a neighbour of i
        int              ncol = 0;
        for(int i = 0; i < m; i++)
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

    int lanes_for(double mean_row)
    {
        int l = 1;
        while(l < 64 && 4 * l < mean_row)
            l <<= 1;
        return l;
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
        s.wg    = m <= MC_WG_ROWS;
        if(s.wg)
        {
            // one pass of the workgroup's lane groups should cover the largest
            // colour, so a colour costs one round of loads, not several
            int widest = 1;
            for(int k = 0; k < ncol; k++)
                widest = std::max(widest, s.off[k + 1] - s.off[k]);
            while(s.lanes > 1 && (1024 / s.lanes) < widest)
                s.lanes >>= 1;
        }
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
        if(s.wg)
        {
            mc_wg_kernel<<<1, 1024, sizeof(double) * m, c->stream>>>(
                m, s.ncol, s.off_d, s.perm, s.prp, s.pci, s.pv, s.pdinv, b, x, s.lanes,
                sym ? 1 : 0);
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
