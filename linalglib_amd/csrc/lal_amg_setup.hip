// SA-AMG setup kernels: strength of connection, PMIS aggregation and the
// smoothed prolongator.
//
// Each decision and each floating-point value is formed by the same
// operations in the same order as the host path -- a lane walking its row in
// CSR order, or a group of lanes whose reductions keep the CSR order (entry
// positions as tie-breaks, folds taken lane by lane in entry order):
//   strength     backend/host/host_amg_strength.cpp:33-70
//   PMIS         backend/host/host_amg_aggregation.cpp:32-221 driven as
//                iterative_solvers/amg/amg_aggregation.cpp:106-215
//   prolongator  iterative_solvers/amg/saamg.cpp:51-207
// The library is built with -ffp-contract=off, and HIP's double sqrt and
// division are correctly rounded, so the results are bit-identical.
#include "lal_internal.h"

#include <algorithm>
#include <climits>

namespace
{
    __global__ void __launch_bounds__(256)
        strength_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                        const double* __restrict__ val, const double* __restrict__ d, double eps,
                        int* __restrict__ conn)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        const double di = fabs(d[i]);
        for(int j = rp[i]; j < rp[i + 1]; j++)
        {
            int    c = ci[j];
            double v = val[j];
            conn[j]  = (c != i) && (fabs(v) >= eps * sqrt(di * fabs(d[c])));
        }
    }

    __device__ __forceinline__ int hash1(unsigned x)
    {
        x = ((x >> 16) ^ x) * 0x45d9f3bu;
        x = ((x >> 16) ^ x) * 0x45d9f3bu;
        x = (x >> 16) ^ x;
        return (int)(x / 2);
    }

    __global__ void __launch_bounds__(256)
        pmis_init_kernel(int m, const int* __restrict__ rp, const int* __restrict__ conn,
                         int* __restrict__ state, int* __restrict__ hash)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        int s = -2;
        for(int j = rp[i]; j < rp[i + 1]; j++)
            if(conn[j] == 1)
            {
                s = 0;
                break;
            }
        state[i] = s;
        hash[i]  = hash1((unsigned)i);
    }

    // Candidate (s, h, r) replaces the running maximum when its state is
    // strictly greater, or equal with a strictly greater hash.
    __device__ __forceinline__ void lex_max(int& ms, int& mh, int& mr, int s, int h, int r)
    {
        if(s > ms || (s == ms && h > mh))
        {
            ms = s;
            mh = h;
            mr = r;
        }
    }

    // PASS: decided rows also copy their state into max_state, so that the
    // setup's own loop can alternate two buffers instead of copying state
    // from max_state before every round (the C ABI round leaves them alone,
    // as the reference's kernel does).
    template <bool PASS>
    __global__ void __launch_bounds__(256)
        pmis_round_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                          const int* __restrict__ conn, const int* __restrict__ state,
                          const int* __restrict__ hash, int64_t* __restrict__ agg,
                          int* __restrict__ max_state, int* __restrict__ incomplete)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        const int si = state[i];
        if(si != 0)
        {
            if(PASS)
                max_state[i] = si;
            return; // only undecided rows change anything
        }
        int ms = si, mh = hash[i], mr = i;
        for(int j = rp[i]; j < rp[i + 1]; j++)
            if(conn[j] == 1)
            {
                int c = ci[j];
                lex_max(ms, mh, mr, state[c], hash[c], c);
            }
        const int r2 = mr;
        for(int j = rp[r2]; j < rp[r2 + 1]; j++)
            if(conn[j] == 1)
            {
                int c = ci[j];
                lex_max(ms, mh, mr, state[c], hash[c], c);
            }
        if(mr == i)
        {
            max_state[i] = 1;
            agg[i]       = 1;
        }
        else if(ms == 1)
        {
            max_state[i] = -1;
            agg[i]       = 0;
        }
        else
            *incomplete = 1;
    }

    __global__ void __launch_bounds__(256)
        pmis_join_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                         const int* __restrict__ conn, const int* __restrict__ state,
                         int64_t* __restrict__ agg, int* __restrict__ max_state)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        const int si = state[i];
        if(si == -1)
        {
            for(int j = rp[i]; j < rp[i + 1]; j++)
                if(conn[j] == 1)
                {
                    int c = ci[j];
                    if(state[c] == 1)
                    {
                        // c has state 1 in the snapshot, so its aggregate is
                        // not written during this pass
                        agg[i]       = agg[c];
                        max_state[i] = 1;
                        break;
                    }
                }
        }
        else if(si == -2)
            agg[i] = -2;
    }

    __device__ __forceinline__ bool p_candidate(int i, int c, int cn)
    {
        return c == i || cn == 1;
    }
}

namespace
{
    // The prolongator P = (I - relax D^-1 A_F) T (saamg.cpp:51-207) with G lanes
    // per row: a wavefront (G = 64) for the long rows of the coarse Galerkin
    // levels, 8 lanes for the finest level's stencil rows.  Row i's entries
    // are its candidates' distinct aggregates >= 0 (the diagonal and the
    // strong neighbours) in ascending order, each the CSR-order fold of its
    // contributions from the first one, with the weak couplings lumped into
    // the diagonal.  Lanes read the row in coalesced chunks of G, the next
    // aggregate in ascending order is a group-wide minimum, and every fold
    // runs over the lanes in entry order.
    template <int G>
    __device__ __forceinline__ int64_t group_min_i64(int64_t v)
    {
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
        {
            const int64_t t = __shfl_xor(v, o, 64);
            v               = t < v ? t : v;
        }
        return v;
    }

    // Smallest aggregate id above `last` among the row's candidates (LLONG_MAX
    // when there is none); group-uniform.
    template <int G>
    __device__ __forceinline__ int64_t next_aggregate(int i, int rs, int re, int gl,
                                                      const int* __restrict__ ci,
                                                      const int* __restrict__ conn,
                                                      const int64_t* __restrict__ agg,
                                                      int64_t last)
    {
        int64_t best = LLONG_MAX;
        for(int j = rs + gl; j < re; j += G)
        {
            const int c = ci[j];
            if(p_candidate(i, c, conn[j]))
            {
                const int64_t a = agg[c];
                if(a > last && a < best)
                    best = a;
            }
        }
        return group_min_i64<G>(best);
    }

    template <int G>
    __global__ void __launch_bounds__(256)
        prolong_count_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                   const int* __restrict__ conn, const int64_t* __restrict__ agg,
                                   int* __restrict__ count)
    {
        const int gl = (int)threadIdx.x % G;
        const int i  = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
        if(i > m)
            return;
        if(i == m)
        {
            if(gl == 0)
                count[m] = 0;
            return;
        }
        const int rs = rp[i], re = rp[i + 1];
        int       k    = 0;
        int64_t   last = -1; // aggregates >= 0 only
        for(;;)
        {
            const int64_t best = next_aggregate<G>(i, rs, re, gl, ci, conn, agg, last);
            if(best == LLONG_MAX)
                break;
            k++;
            last = best;
        }
        if(gl == 0)
            count[i] = k;
    }

    template <int G>
    __global__ void __launch_bounds__(256)
        prolong_fill_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                  const double* __restrict__ val, const int* __restrict__ conn,
                                  const int64_t* __restrict__ agg, double relax,
                                  const int* __restrict__ prp, int* __restrict__ pci,
                                  double* __restrict__ pv)
    {
        const int gl = (int)threadIdx.x % G;
        const int gb = ((int)threadIdx.x & 63) - gl; // the group's first lane
        const int i  = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
        if(i >= m)
            return;
        const int rs = rp[i], re = rp[i + 1];
        // a_ii minus the weak couplings, summed in CSR order
        double diagonal = 0.0;
        for(int j0 = rs; j0 < re; j0 += G)
        {
            const int j    = j0 + gl;
            int       kind = 0;
            double    v    = 0.0;
            if(j < re)
            {
                const int c = ci[j];
                if(c == i)
                    kind = 1;
                else if(conn[j] == 0)
                    kind = 2;
                v = val[j];
            }
            const int n = min(G, re - j0);
            for(int t = 0; t < n; t++)
            {
                const int    kt = __shfl(kind, gb + t, 64);
                const double vt = __shfl(v, gb + t, 64);
                if(kt == 1)
                    diagonal += vt;
                else if(kt == 2)
                    diagonal -= vt;
            }
        }
        const double invd = 1.0 / diagonal;
        int64_t      last = -1;
        for(int out = prp[i]; out < prp[i + 1]; out++)
        {
            const int64_t best  = next_aggregate<G>(i, rs, re, gl, ci, conn, agg, last);
            bool          first = true;
            double        acc   = 0.0;
            for(int j0 = rs; j0 < re; j0 += G)
            {
                const int j     = j0 + gl;
                bool      match = false;
                double    value = 0.0;
                if(j < re)
                {
                    const int c = ci[j];
                    if(p_candidate(i, c, conn[j]) && agg[c] == best)
                    {
                        match = true;
                        value = (c == i) ? 1.0 - relax : -relax * invd * val[j];
                    }
                }
                unsigned long long mask = __ballot(match) >> gb;
                if(G < 64)
                    mask &= (1ull << G) - 1ull;
                while(mask)
                {
                    const int    t  = __ffsll((long long)mask) - 1;
                    const double vt = __shfl(value, gb + t, 64);
                    mask &= mask - 1;
                    if(first)
                    {
                        acc   = vt;
                        first = false;
                    }
                    else
                        acc += vt;
                }
            }
            if(gl == 0)
            {
                pci[out] = (int)best;
                pv[out]  = acc;
            }
            last = best;
        }
    }

    // Rows this long on average take a wavefront per row, shorter ones 8 lanes.
    constexpr long PROLONG_WAVE_MIN_ROW = 24;
}

namespace
{
    // Strength, the diagonal and the PMIS passes with G lanes per row (G = 16
    // or 64, 256 / G rows per workgroup) for the long rows of the coarse
    // Galerkin levels, where a lane walking a row of 40-190 entries alone is
    // a chain of dependent loads.  Each lane takes the row's entries
    // j = rs + gl, rs + gl + G, ... in order; every per-row result the
    // thread-per-row kernels form in CSR order -- the first (or last) match,
    // a lexicographic maximum that keeps its first occurrence -- is the
    // group's reduction with the entry position as the tie-break, so the
    // results are bit for bit the same.
    template <int G>
    __device__ __forceinline__ int group_row(int& gl)
    {
        gl = (int)threadIdx.x % G;
        return blockIdx.x * (256 / G) + (int)threadIdx.x / G;
    }

    template <int G>
    __device__ __forceinline__ int group_min(int v)
    {
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
            v = min(v, __shfl_xor(v, o, 64));
        return v;
    }

    template <int G>
    __device__ __forceinline__ int group_max(int v)
    {
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
            v = max(v, __shfl_xor(v, o, 64));
        return v;
    }

    // d[i] = the row's first diagonal entry (0 when it has none)
    template <int G>
    __global__ void __launch_bounds__(256)
        diagonal_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ val, double* __restrict__ d)
    {
        int       gl;
        const int i = group_row<G>(gl);
        if(i >= m)
            return;
        const int rs = rp[i], re = rp[i + 1];
        int       first = INT_MAX;
        for(int j = rs + gl; j < re; j += G)
            if(ci[j] == i)
            {
                first = j;
                break;
            }
        first = group_min<G>(first);
        if(gl == 0 && first != INT_MAX)
            d[i] = val[first];
    }

    template <int G>
    __global__ void __launch_bounds__(256)
        strength_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                              const double* __restrict__ val, const double* __restrict__ d,
                              double eps, int* __restrict__ conn)
    {
        int       gl;
        const int i = group_row<G>(gl);
        if(i >= m)
            return;
        const double di = fabs(d[i]);
        for(int j = rp[i] + gl; j < rp[i + 1]; j += G)
        {
            int    c = ci[j];
            double v = val[j];
            conn[j]  = (c != i) && (fabs(v) >= eps * sqrt(di * fabs(d[c])));
        }
    }

    template <int G>
    __global__ void __launch_bounds__(256)
        pmis_init_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ conn,
                               int* __restrict__ state, int* __restrict__ hash)
    {
        int       gl;
        const int i = group_row<G>(gl);
        if(i >= m)
            return;
        int strong = 0;
        for(int j = rp[i] + gl; j < rp[i + 1] && !strong; j += G)
            strong = conn[j] == 1;
        strong = group_max<G>(strong);
        if(gl == 0)
        {
            state[i] = strong ? 0 : -2;
            hash[i]  = hash1((unsigned)i);
        }
    }

    // A PMIS candidate: state, hash, node, and its position in the scan.
    struct pmis_cand
    {
        int s, h, r, pos;
    };

    // b replaces a when (s, h) is lexicographically greater, or equal and
    // earlier in the scan.
    __device__ __forceinline__ void pmis_take(pmis_cand& a, const pmis_cand& b)
    {
        if(b.s > a.s || (b.s == a.s && (b.h > a.h || (b.h == a.h && b.pos < a.pos))))
            a = b;
    }

    // The best strong neighbour of row `row` across the group: each lane keeps
    // the first maximum of its entries, the group the earliest of the lanes'.
    template <int G>
    __device__ __forceinline__ pmis_cand pmis_scan(int row, int gl, const int* __restrict__ rp,
                                                   const int* __restrict__ ci,
                                                   const int* __restrict__ conn,
                                                   const int* __restrict__ state,
                                                   const int* __restrict__ hash)
    {
        pmis_cand best{INT_MIN, INT_MIN, -1, INT_MAX};
        const int rs = rp[row], re = rp[row + 1];
        for(int j = rs + gl; j < re; j += G)
            if(conn[j] == 1)
            {
                const int c = ci[j];
                const int sc = state[c], hc = hash[c];
                if(sc > best.s || (sc == best.s && hc > best.h))
                    best = pmis_cand{sc, hc, c, j - rs};
            }
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
        {
            const pmis_cand other{__shfl_xor(best.s, o, 64), __shfl_xor(best.h, o, 64),
                                  __shfl_xor(best.r, o, 64), __shfl_xor(best.pos, o, 64)};
            pmis_take(best, other);
        }
        return best;
    }

    // pmis_round_kernel with G lanes per row: the running maximum starts at
    // the row itself, takes the row's strong neighbours in order, then the
    // strong neighbours of the row that maximum names.
    template <int G, bool PASS>
    __global__ void __launch_bounds__(256)
        pmis_round_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                const int* __restrict__ conn, const int* __restrict__ state,
                                const int* __restrict__ hash, int64_t* __restrict__ agg,
                                int* __restrict__ max_state, int* __restrict__ incomplete)
    {
        int       gl;
        const int i = group_row<G>(gl);
        if(i >= m)
            return;
        const int si = state[i];
        if(si != 0)
        {
            if(PASS && gl == 0)
                max_state[i] = si;
            return; // only undecided rows change anything
        }
        // the row itself precedes its neighbours, so it wins a tie
        pmis_cand run{si, hash[i], i, -1};
        pmis_cand c1 = pmis_scan<G>(i, gl, rp, ci, conn, state, hash);
        if(c1.s > run.s || (c1.s == run.s && c1.h > run.h))
            run = c1;
        // the second row's entries all follow the first scan's
        pmis_cand c2 = pmis_scan<G>(run.r, gl, rp, ci, conn, state, hash);
        if(c2.s > run.s || (c2.s == run.s && c2.h > run.h))
            run = c2;
        if(gl != 0)
            return;
        if(run.r == i)
        {
            max_state[i] = 1;
            agg[i]       = 1;
        }
        else if(run.s == 1)
        {
            max_state[i] = -1;
            agg[i]       = 0;
        }
        else
            *incomplete = 1;
    }

    template <int G>
    __global__ void __launch_bounds__(256)
        pmis_join_group_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                               const int* __restrict__ conn, const int* __restrict__ state,
                               int64_t* __restrict__ agg, int* __restrict__ max_state)
    {
        int       gl;
        const int i = group_row<G>(gl);
        if(i >= m)
            return;
        const int si = state[i];
        if(si == -1)
        {
            // the first strong neighbour (in CSR order) with state 1
            int first = INT_MAX;
            for(int j = rp[i] + gl; j < rp[i + 1]; j += G)
                if(conn[j] == 1 && state[ci[j]] == 1)
                {
                    first = j;
                    break;
                }
            first = group_min<G>(first);
            if(gl == 0 && first != INT_MAX)
            {
                // c has state 1 in the snapshot, so its aggregate is not
                // written during this pass
                agg[i]       = agg[ci[first]];
                max_state[i] = 1;
            }
        }
        else if(si == -2 && gl == 0)
            agg[i] = -2;
    }

    // Lanes per row for the row-walking setup passes: one for short rows (the
    // finest level's stencil), 16 or 64 for the coarse Galerkin levels.
    inline int setup_row_lanes(long nnz, int m)
    {
        const long avg = m > 0 ? nnz / m : 0;
        return avg >= 48 ? 64 : avg >= 12 ? 16 : 1;
    }

    inline int group_grid(int m, int lanes)
    {
        return (int)(((long)m * lanes + 255) / 256);
    }
}

namespace lal
{
    int diagonal_f64(lal_context* c, int m, const int* rp, const int* ci, const double* v,
                     double* d);

    int sa_strength(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                    const double* v, double eps, int* conn)
    {
        if(m == 0)
            return LAL_OK;
        double* d = nullptr;
        LAL_HIP(hipMallocAsync((void**)&d, sizeof(double) * m, c->stream));
        LAL_HIP(hipMemsetAsync(d, 0, sizeof(double) * m, c->stream));
        const int lanes = setup_row_lanes(nnz, m);
        if(lanes == 64)
        {
            diagonal_group_kernel<64><<<group_grid(m, 64), 256, 0, c->stream>>>(m, rp, ci, v, d);
            LAL_LAUNCHED();
            strength_group_kernel<64>
                <<<group_grid(m, 64), 256, 0, c->stream>>>(m, rp, ci, v, d, eps, conn);
        }
        else if(lanes == 16)
        {
            diagonal_group_kernel<16><<<group_grid(m, 16), 256, 0, c->stream>>>(m, rp, ci, v, d);
            LAL_LAUNCHED();
            strength_group_kernel<16>
                <<<group_grid(m, 16), 256, 0, c->stream>>>(m, rp, ci, v, d, eps, conn);
        }
        else
        {
            LAL_TRY(diagonal_f64(c, m, rp, ci, v, d));
            strength_kernel<<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, v, d, eps, conn);
        }
        LAL_LAUNCHED();
        LAL_HIP(hipFreeAsync(d, c->stream));
        return LAL_OK;
    }

    int pmis_aggregate(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                       const int* conn, int64_t* agg)
    {
        if(m == 0)
            return LAL_OK;
        int *state = nullptr, *max_state = nullptr, *hash = nullptr;
        LAL_HIP(hipMallocAsync((void**)&state, sizeof(int) * m, c->stream));
        LAL_HIP(hipMallocAsync((void**)&max_state, sizeof(int) * m, c->stream));
        LAL_HIP(hipMallocAsync((void**)&hash, sizeof(int) * m, c->stream));
        LAL_HIP(hipMemsetAsync(agg, 0, sizeof(int64_t) * m, c->stream));
        const int g     = grid_for(m, 256);
        const int lanes = setup_row_lanes(nnz, m);
        const int gg    = group_grid(m, lanes);
        if(lanes == 64)
            pmis_init_group_kernel<64><<<gg, 256, 0, c->stream>>>(m, rp, conn, max_state, hash);
        else if(lanes == 16)
            pmis_init_group_kernel<16><<<gg, 256, 0, c->stream>>>(m, rp, conn, max_state, hash);
        else
            pmis_init_kernel<<<g, 256, 0, c->stream>>>(m, rp, conn, max_state, hash);
        LAL_LAUNCHED();
        // The reference's loop (amg_aggregation.cpp:106-215): up to 20 rounds,
        // state = max_state before each, stop after the first complete one.
        // Here the rounds alternate between two buffers (the round kernel
        // writes every decided row through, and a row still undecided was
        // undecided in the older buffer too, so it reads 0 there as in the
        // copy), and the host checks after every second round: a round after
        // a complete one finds no undecided row and changes nothing, so the
        // aggregates are those of the reference's loop.
        LAL_HIP(hipMemcpyAsync(state, max_state, sizeof(int) * m, hipMemcpyDeviceToDevice,
                               c->stream));
        int* buf[2] = {max_state, state}; // buf[cur]: the latest state
        int  cur    = 0;
        auto round  = [&](int* flag) -> int {
            const int* in  = buf[cur];
            int*       out = buf[cur ^ 1];
            if(lanes == 64)
                pmis_round_group_kernel<64, true><<<gg, 256, 0, c->stream>>>(
                    m, rp, ci, conn, in, hash, agg, out, flag);
            else if(lanes == 16)
                pmis_round_group_kernel<16, true><<<gg, 256, 0, c->stream>>>(
                    m, rp, ci, conn, in, hash, agg, out, flag);
            else
                pmis_round_kernel<true><<<g, 256, 0, c->stream>>>(m, rp, ci, conn, in, hash, agg,
                                                                  out, flag);
            LAL_LAUNCHED();
            cur ^= 1;
            return LAL_OK;
        };
        for(int iter = 0; iter < 20; iter += 2)
        {
            LAL_HIP(hipMemsetAsync(c->d_flag, 0, 2 * sizeof(int), c->stream));
            LAL_TRY(round(c->d_flag));
            LAL_TRY(round(c->d_flag + 1));
            LAL_HIP(hipMemcpyAsync(c->h_flag, c->d_flag + 1, sizeof(int), hipMemcpyDeviceToHost,
                                   c->stream));
            LAL_HIP(hipStreamSynchronize(c->stream));
            if(c->h_flag[0] == 0)
                break;
        }
        // the joins below read max_state: the latest state must be there
        if(buf[cur] != max_state)
            LAL_HIP(hipMemcpyAsync(max_state, buf[cur], sizeof(int) * m, hipMemcpyDeviceToDevice,
                                   c->stream));
        LAL_TRY(scan_i64_dev(c, m, agg));
        for(int k = 0; k < 2; k++)
        {
            LAL_HIP(hipMemcpyAsync(state, max_state, sizeof(int) * m, hipMemcpyDeviceToDevice,
                                   c->stream));
            if(lanes == 64)
                pmis_join_group_kernel<64>
                    <<<gg, 256, 0, c->stream>>>(m, rp, ci, conn, state, agg, max_state);
            else if(lanes == 16)
                pmis_join_group_kernel<16>
                    <<<gg, 256, 0, c->stream>>>(m, rp, ci, conn, state, agg, max_state);
            else
                pmis_join_kernel<<<g, 256, 0, c->stream>>>(m, rp, ci, conn, state, agg,
                                                           max_state);
            LAL_LAUNCHED();
        }
        LAL_HIP(hipFreeAsync(state, c->stream));
        LAL_HIP(hipFreeAsync(max_state, c->stream));
        LAL_HIP(hipFreeAsync(hash, c->stream));
        return LAL_OK;
    }

    int sa_prolongator_nnz(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                           const int* conn, const int64_t* agg, int* prp, int* nnz_P,
                           int* n_c)
    {
        // count[m] = 0 is written by the group of row m, so m + 1 groups
        if(m > 0 && nnz >= PROLONG_WAVE_MIN_ROW * (long)m)
            prolong_count_group_kernel<64><<<(m + 1 + 3) / 4, 256, 0, c->stream>>>(m, rp, ci,
                                                                                   conn, agg, prp);
        else
            prolong_count_group_kernel<8><<<(m + 1 + 31) / 32, 256, 0, c->stream>>>(m, rp, ci,
                                                                                    conn, agg, prp);
        LAL_LAUNCHED();
        LAL_TRY(scan_i32_dev(c, m + 1, prp));
        int64_t mx = LLONG_MIN;
        LAL_TRY(max_i64_dev(c, m, agg, &mx));
        int64_t nc = std::max<int64_t>(mx, -1) + 1;
        if(nc > INT_MAX)
            return LAL_ERR_UNSUPPORTED;
        LAL_HIP(hipMemcpyAsync(nnz_P, prp + m, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        *n_c = (int)nc;
        return LAL_OK;
    }

    int sa_prolongator_fill(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                            const double* v, const int* conn, const int64_t* agg, double relax,
                            const int* prp, int* pci, double* pv)
    {
        if(m == 0)
            return LAL_OK;
        if(nnz >= PROLONG_WAVE_MIN_ROW * (long)m)
            prolong_fill_group_kernel<64><<<(m + 3) / 4, 256, 0, c->stream>>>(
                m, rp, ci, v, conn, agg, relax, prp, pci, pv);
        else
            prolong_fill_group_kernel<8><<<(m + 31) / 32, 256, 0, c->stream>>>(
                m, rp, ci, v, conn, agg, relax, prp, pci, pv);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

namespace
{
    // Unsmoothed aggregation (uaamg.cpp:48-123): row i of P holds one entry,
    // 1.0 in column agg[i], when the row was aggregated and nothing otherwise.
    __global__ void __launch_bounds__(256)
        ua_count_kernel(int m, const int64_t* __restrict__ agg, int* __restrict__ prp)
    {
        const int i = blockIdx.x * 256 + threadIdx.x;
        if(i <= m)
            prp[i] = (i < m && agg[i] >= 0) ? 1 : 0;
    }

    __global__ void __launch_bounds__(256)
        ua_fill_kernel(int m, const int64_t* __restrict__ agg, const int* __restrict__ prp,
                       int* __restrict__ pci, double* __restrict__ pv)
    {
        const int i = blockIdx.x * 256 + threadIdx.x;
        if(i < m && agg[i] >= 0)
        {
            pci[prp[i]] = (int)agg[i];
            pv[prp[i]]  = 1.0;
        }
    }
}

namespace lal
{
    int ua_prolongator_nnz(lal_context* c, int m, const int64_t* agg, int* prp, int* nnz_P,
                           int* n_c)
    {
        ua_count_kernel<<<grid_for(m + 1, 256), 256, 0, c->stream>>>(m, agg, prp);
        LAL_LAUNCHED();
        LAL_TRY(scan_i32_dev(c, m + 1, prp));
        int64_t mx = LLONG_MIN;
        LAL_TRY(max_i64_dev(c, m, agg, &mx));
        const int64_t nc = std::max<int64_t>(mx, -1) + 1;
        if(nc > INT_MAX)
            return LAL_ERR_UNSUPPORTED;
        LAL_HIP(hipMemcpyAsync(nnz_P, prp + m, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        *n_c = (int)nc;
        return LAL_OK;
    }

    int ua_prolongator_fill(lal_context* c, int m, const int64_t* agg, const int* prp, int* pci,
                            double* pv)
    {
        if(m == 0)
            return LAL_OK;
        ua_fill_kernel<<<grid_for(m, 256), 256, 0, c->stream>>>(m, agg, prp, pci, pv);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

extern "C" int lal_sa_strength_f64(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                                   const double* v, double eps, int* conn)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && (nnz == 0 || (ci && v && conn))));
    return lal::sa_strength(c, m, nnz, rp, ci, v, eps, conn);
}

extern "C" int lal_pmis_aggregate(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                                  const int* conn, int64_t* agg)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && agg && (nnz == 0 || (ci && conn))));
    return lal::pmis_aggregate(c, m, nnz, rp, ci, conn, agg);
}

extern "C" int lal_pmis_init(lal_handle c, int m, int nnz, const int* rp, const int* conn,
                             int* state, int* hash)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && state && hash && (nnz == 0 || conn)));
    if(m == 0)
        return LAL_OK;
    pmis_init_kernel<<<lal::grid_for(m, 256), 256, 0, c->stream>>>(m, rp, conn, state, hash);
    LAL_LAUNCHED();
    return LAL_OK;
}

extern "C" int lal_pmis_round(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                              const int* conn, const int* state, const int* hash, int64_t* agg,
                              int* max_state, int* complete)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0 && complete != nullptr);
    LAL_REQUIRE(m == 0 || (rp && state && hash && agg && max_state && (nnz == 0 || (ci && conn))));
    *complete = 1;
    if(m == 0)
        return LAL_OK;
    LAL_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
    pmis_round_kernel<false><<<lal::grid_for(m, 256), 256, 0, c->stream>>>(
        m, rp, ci, conn, state, hash, agg, max_state, c->d_flag);
    LAL_LAUNCHED();
    LAL_HIP(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LAL_HIP(hipStreamSynchronize(c->stream));
    *complete = c->h_flag[0] == 0;
    return LAL_OK;
}

extern "C" int lal_pmis_join(lal_handle c, int m, int nnz, const int* rp, const int* ci,
                             const int* conn, const int* state, int64_t* agg, int* max_state)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && state && agg && max_state && (nnz == 0 || (ci && conn))));
    if(m == 0)
        return LAL_OK;
    pmis_join_kernel<<<lal::grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, conn, state, agg,
                                                                   max_state);
    LAL_LAUNCHED();
    return LAL_OK;
}

extern "C" int lal_sa_prolongator_nnz(lal_handle c, int m, int nnz, const int* rp,
                                      const int* ci, const int* conn, const int64_t* agg,
                                      int* prp, int* nnz_P, int* n_c)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0 && prp && nnz_P && n_c);
    LAL_REQUIRE(m == 0 || (rp && agg && (nnz == 0 || (ci && conn))));
    return lal::sa_prolongator_nnz(c, m, nnz, rp, ci, conn, agg, prp, nnz_P, n_c);
}

extern "C" int lal_sa_prolongator_fill_f64(lal_handle c, int m, int nnz, const int* rp,
                                           const int* ci, const double* v, const int* conn,
                                           const int64_t* agg, double relax, const int* prp,
                                           int* pci, double* pv)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && nnz >= 0);
    LAL_REQUIRE(m == 0 || (rp && agg && prp && (nnz == 0 || (ci && v && conn))));
    return lal::sa_prolongator_fill(c, m, nnz, rp, ci, v, conn, agg, relax, prp, pci, pv);
}
