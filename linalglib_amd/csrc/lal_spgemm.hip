// Sparse products of the SA-AMG setup: transpose (R = P^T) and the two-pass
// hash SpGEMM used for the Galerkin product A_c = R (A P).
//
// Transpose: count, scan, scatter, then each row of the result is sorted by
// column, which gives the stable order of backend/host/host_csr2csc.cpp:30-75.
//
// SpGEMM: rows of C are binned on the device by their product count into LDS
// hash tables of H = 64 ... 4096 slots (open addressing, linear probing,
// slots claimed with atomicCAS); rows beyond that get a table in global
// memory.  Tables of up to 256 slots are built by H / 8 lanes per row (and
// the symbolic pass's key-only tables of 512 and 1024 slots by 32), larger
// ones by a wavefront per row.  The numeric pass walks the A row
// sequentially and lets the lanes take the B row in parallel; the columns of one B row are distinct, so
// every C entry receives at most one contribution per step and is folded in
// A-row-then-B-row order from its first product -- the exact accumulation of
// backend/host/host_csrgemm.cpp:91-183.  The table is then compacted and each
// entry is placed by its rank, giving ascending columns.
#include "lal_internal.h"

#include <algorithm>
#include <vector>

namespace
{
    // ------------------------------------------------------------ transpose
    __global__ void __launch_bounds__(256)
        col_count_kernel(int nnz, const int* __restrict__ ci, int* __restrict__ cnt)
    {
        int k = blockIdx.x * 256 + threadIdx.x;
        if(k < nnz)
            atomicAdd(&cnt[ci[k]], 1);
    }

    template <typename T>
    __global__ void __launch_bounds__(256)
        transpose_scatter_kernel(int m, const int* __restrict__ rp, const int* __restrict__ ci,
                                 const T* __restrict__ v, int* __restrict__ cursor,
                                 int* __restrict__ tci, T* __restrict__ tv)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        for(int j = rp[i]; j < rp[i + 1]; j++)
        {
            int pos  = atomicAdd(&cursor[ci[j]], 1);
            tci[pos] = i;
            tv[pos]  = v[j];
        }
    }

    // Sorts each row by column, stably, with one wavefront per row: each lane
    // ranks its entries (up to four) by counting the row's smaller columns,
    // ties broken by position; every load of the row completes before the
    // first store, so the row is permuted in place.  Rows longer than 256
    // entries take lane 0's insertion sort.
    template <typename T>
    __global__ void __launch_bounds__(256)
        row_sort_wave_kernel(int m, const int* __restrict__ rp, int* __restrict__ ci,
                             T* __restrict__ v)
    {
        const int lane = threadIdx.x & 63;
        const int i    = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
        if(i >= m)
            return;
        const int s = rp[i], e = rp[i + 1];
        if(e - s <= 1)
            return;
        if(e - s > 256)
        {
            if(lane != 0)
                return;
            for(int j = s + 1; j < e; j++)
            {
                int kc = ci[j];
                T   kv = v[j];
                int    t  = j - 1;
                while(t >= s && ci[t] > kc)
                {
                    ci[t + 1] = ci[t];
                    v[t + 1]  = v[t];
                    t--;
                }
                ci[t + 1] = kc;
                v[t + 1]  = kv;
            }
            return;
        }
        int key[4], rank[4];
        T   val[4];
#pragma unroll
        for(int u = 0; u < 4; u++)
        {
            const int j = s + lane + 64 * u;
            key[u]      = j < e ? ci[j] : INT_MAX;
            val[u]      = j < e ? v[j] : (T)0;
            rank[u]     = 0;
        }
        for(int t = s; t < e; t++)
        {
            const int kt = ci[t];
#pragma unroll
            for(int u = 0; u < 4; u++)
                rank[u] += kt < key[u] || (kt == key[u] && t < s + lane + 64 * u);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for(int u = 0; u < 4; u++)
            if(s + lane + 64 * u < e)
            {
                ci[s + rank[u]] = key[u];
                v[s + rank[u]]  = val[u];
            }
    }

    // -------------------------------------------------------------- SpGEMM
    __global__ void __launch_bounds__(256)
        products_kernel(int m, const int* __restrict__ arp, const int* __restrict__ aci,
                        const int* __restrict__ brp, int* __restrict__ prod)
    {
        int i = blockIdx.x * 256 + threadIdx.x;
        if(i >= m)
            return;
        long long p = 0;
        for(int j = arp[i]; j < arp[i + 1]; j++)
        {
            int c = aci[j];
            p += brp[c + 1] - brp[c];
        }
        prod[i] = (int)min(p, (long long)INT_MAX);
    }

    // products_kernel with G lanes per row, for long rows of A: the lanes
    // sum strided parts of the row, the group adds them up (integers: the
    // same count).
    template <int G>
    __global__ void __launch_bounds__(256)
        products_group_kernel(int m, const int* __restrict__ arp, const int* __restrict__ aci,
                              const int* __restrict__ brp, int* __restrict__ prod)
    {
        const int gl = (int)threadIdx.x % G;
        const int i  = blockIdx.x * (256 / G) + (int)threadIdx.x / G;
        if(i >= m)
            return;
        long long p = 0;
        for(int j = arp[i] + gl; j < arp[i + 1]; j += G)
        {
            int c = aci[j];
            p += brp[c + 1] - brp[c];
        }
#pragma unroll
        for(int o = G / 2; o > 0; o >>= 1)
            p += __shfl_xor(p, o, 64);
        if(gl == 0)
            prod[i] = (int)min(p, (long long)INT_MAX);
    }

    __device__ __forceinline__ int slot_of(int col, int hmask)
    {
        return (int)(((unsigned)col * 0x9E3779B1u) >> 7) & hmask;
    }

    // Orders one step's table updates before the next step's reads across
    // the lanes of the wave: LDS tables need workgroup scope, tables in global
    // memory agent scope (the L1 is not coherent with the atomics in L2).
    template <bool GLOBAL>
    __device__ __forceinline__ void wave_sync_mem()
    {
        if(GLOBAL)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        else
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        if(GLOBAL)
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        else
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }

    // One wavefront builds row `row` of C in a table of H slots (H a power of
    // two, at least 64 and at least twice the row's distinct columns).
    //
    // The numeric pass sizes the table by the row's nonzero count from the
    // symbolic pass.  Should a caller pass a row pointer that undercounts a
    // row, the probing stops after H slots and the row is cut to its length,
    // *err is set, and no slot or output entry outside the row is touched.
    template <bool NUMERIC, bool GLOBAL>
    __device__ void spgemm_row(int row, int* keys, double* vals, int H, int lane,
                               const int* __restrict__ arp, const int* __restrict__ aci,
                               const double* __restrict__ av, const int* __restrict__ brp,
                               const int* __restrict__ bci, const double* __restrict__ bv,
                               const int* __restrict__ crp, int* __restrict__ cout_,
                               double* __restrict__ cval, int* __restrict__ err)
    {
        const int hmask = H - 1;
        for(int s = lane; s < H; s += 64)
            keys[s] = -1;
        wave_sync_mem<GLOBAL>();
        int local = 0;
        for(int j = arp[row]; j < arp[row + 1]; j++)
        {
            const int    ca = aci[j];
            const double va = NUMERIC ? 1.0 * av[j] : 0.0;
            const int    be = brp[ca + 1];
            for(int p = brp[ca] + lane; p < be; p += 64)
            {
                const int col = bci[p];
                int       h   = slot_of(col, hmask);
                for(int probe = 0;; probe++)
                {
                    if(probe == H)
                    {
                        *err = 1; // table full: the row pointer undercounts this row
                        break;
                    }
                    int old = atomicCAS(&keys[h], -1, col);
                    if(old == -1)
                    {
                        local++;
                        if(NUMERIC)
                            vals[h] = va * bv[p];
                        break;
                    }
                    if(old == col)
                    {
                        if(NUMERIC)
                            vals[h] += va * bv[p];
                        break;
                    }
                    h = (h + 1) & hmask;
                }
            }
            wave_sync_mem<GLOBAL>();
        }
        if(!NUMERIC)
        {
            int total = lal::wave_sum(local);
            if(lane == 0)
                cout_[row] = total;
            return;
        }
        // compact the occupied slots into [0, nnz) in place
        int                      base  = 0;
        const unsigned long long below = (1ull << lane) - 1ull;
        for(int s0 = 0; s0 < H; s0 += 64)
        {
            const int          s     = s0 + lane;
            const int          k     = keys[s];
            const double       v     = vals[s];
            const bool         valid = k != -1;
            unsigned long long mask  = __ballot(valid);
            wave_sync_mem<GLOBAL>();
            if(valid)
            {
                int pos   = base + __popcll(mask & below);
                keys[pos] = k;
                vals[pos] = v;
            }
            base += __popcll(mask);
            wave_sync_mem<GLOBAL>();
        }
        const int nnz = base;
        const int out = crp[row];
        const int len = crp[row + 1] - out;
        if(nnz != len && lane == 0)
            *err = 1;
        for(int e = lane; e < nnz; e += 64)
        {
            const int k = keys[e];
            int       r = 0;
            for(int t = 0; t < nnz; t++)
                r += keys[t] < k;
            if(r < len)
            {
                cout_[out + r] = k;
                cval[out + r]  = vals[e];
            }
        }
    }

    // Rows of C in LDS tables of H slots: G lanes per row, NT / G rows per
    // workgroup (G = H / 8 up to 256 slots, 32 for the symbolic pass's 512-
    // and 1024-slot key-only tables, a wavefront otherwise).  The steps and
    // the accumulation order are those of spgemm_row -- A-row entries one
    // after another, the lanes of the group over the entries of that entry's
    // B row -- but the group first loads up to G entries of the A row and
    // their B-row bounds at once, one entry per lane, and each step takes its
    // bounds from the owning lane, so a step waits on one load (the B row)
    // instead of three dependent ones; each lane's first entry of the next
    // step's B row is loaded during the current step.  Rows of P and A P hold
    // a few to a few tens of entries, which a full wavefront per row would
    // leave mostly idle.
    template <int H, int G, int NT, bool NUMERIC>
    __global__ void __launch_bounds__(NT)
        spgemm_group_kernel(int nrows, const int* __restrict__ rows, const int* __restrict__ arp,
                            const int* __restrict__ aci, const double* __restrict__ av,
                            const int* __restrict__ brp, const int* __restrict__ bci,
                            const double* __restrict__ bv, const int* __restrict__ crp,
                            int* __restrict__ cout_, double* __restrict__ cval,
                            int* __restrict__ err)
    {
        static_assert(G >= 2 && G <= 64 && (G & (G - 1)) == 0 && H % G == 0 && NT % G == 0);
        constexpr int RPB   = NT / G;
        constexpr unsigned long long GMASK = G == 64 ? ~0ull : (1ull << (G % 64)) - 1ull;
        constexpr int hmask = H - 1;
        __shared__ int    s_keys[RPB][H];
        __shared__ double s_vals[NUMERIC ? RPB : 1][NUMERIC ? H : 1];
        const int         grp = (int)threadIdx.x / G;
        const int         gl  = (int)threadIdx.x % G;       // lane in the group
        const int         gb  = ((int)threadIdx.x & 63) - gl; // the group's first lane
        const int         r   = blockIdx.x * RPB + grp;
        if(r >= nrows)
            return;
        const int row  = rows[r];
        int*      keys = s_keys[grp];
        double*   vals = NUMERIC ? s_vals[grp] : nullptr;
        for(int s = gl; s < H; s += G)
            keys[s] = -1;
        wave_sync_mem<false>();
        int       local = 0;
        const int rs = arp[row], re = arp[row + 1];
        for(int j0 = rs; j0 < re; j0 += G)
        {
            const int j  = j0 + gl;
            int       bs = 0, be = 0;
            double    va = 0.0;
            if(j < re)
            {
                const int ca = aci[j];
                if(NUMERIC)
                    va = 1.0 * av[j];
                bs = brp[ca];
                be = brp[ca + 1];
            }
            const int cnt = min(G, re - j0);
            // each lane's first entry of step t's B row is loaded during step
            // t - 1, so the step's inserts do not wait on it
            int    nbs = __shfl(bs, gb, 64), nbe = __shfl(be, gb, 64);
            double nva  = NUMERIC ? __shfl(va, gb, 64) : 0.0;
            int    ncol = nbs + gl < nbe ? bci[nbs + gl] : -1;
            double nbv  = NUMERIC && nbs + gl < nbe ? bv[nbs + gl] : 0.0;
            for(int t = 0; t < cnt; t++)
            {
                const int    tbs = nbs, tbe = nbe;
                const double tva = nva;
                const int    col0 = ncol;
                const double bv0  = nbv;
                if(t + 1 < cnt)
                {
                    nbs = __shfl(bs, gb + t + 1, 64);
                    nbe = __shfl(be, gb + t + 1, 64);
                    if(NUMERIC)
                        nva = __shfl(va, gb + t + 1, 64);
                    ncol = nbs + gl < nbe ? bci[nbs + gl] : -1;
                    if(NUMERIC)
                        nbv = nbs + gl < nbe ? bv[nbs + gl] : 0.0;
                }
                for(int p = tbs + gl; p < tbe; p += G)
                {
                    const bool   first = p == tbs + gl;
                    const int    col   = first ? col0 : bci[p];
                    const double bp    = NUMERIC ? (first ? bv0 : bv[p]) : 0.0;
                    int          h     = slot_of(col, hmask);
                    for(int probe = 0;; probe++)
                    {
                        if(probe == H)
                        {
                            *err = 1; // table full: the row pointer undercounts this row
                            break;
                        }
                        int old = atomicCAS(&keys[h], -1, col);
                        if(old == -1)
                        {
                            local++;
                            if(NUMERIC)
                                vals[h] = tva * bp;
                            break;
                        }
                        if(old == col)
                        {
                            if(NUMERIC)
                                vals[h] += tva * bp;
                            break;
                        }
                        h = (h + 1) & hmask;
                    }
                }
                wave_sync_mem<false>();
            }
        }
        if(!NUMERIC)
        {
#pragma unroll
            for(int o = G / 2; o > 0; o >>= 1)
                local += __shfl_xor(local, o, 64);
            if(gl == 0)
                cout_[row] = local;
            return;
        }
        // compact the occupied slots into [0, nnz) in place
        int                      base  = 0;
        const unsigned long long below = (1ull << gl) - 1ull;
        for(int s0 = 0; s0 < H; s0 += G)
        {
            const int                s     = s0 + gl;
            const int                k     = keys[s];
            const double             v     = vals[s];
            const bool               valid = k != -1;
            const unsigned long long mask  = (__ballot(valid) >> gb) & GMASK;
            wave_sync_mem<false>();
            if(valid)
            {
                int pos   = base + __popcll(mask & below);
                keys[pos] = k;
                vals[pos] = v;
            }
            base += __popcll(mask);
            wave_sync_mem<false>();
        }
        const int nnz = base;
        const int out = crp[row];
        const int len = crp[row + 1] - out;
        if(nnz != len && gl == 0)
            *err = 1;
        for(int e = gl; e < nnz; e += G)
        {
            const int k  = keys[e];
            int       rk = 0;
            for(int t = 0; t < nnz; t++)
                rk += keys[t] < k;
            if(rk < len)
            {
                cout_[out + rk] = k;
                cval[out + rk]  = vals[e];
            }
        }
    }

    template <bool NUMERIC>
    __global__ void __launch_bounds__(64)
        spgemm_global_kernel(int nrows, const int* __restrict__ rows,
                             const long long* __restrict__ toff, int* __restrict__ tkeys,
                             double* __restrict__ tvals, const int* __restrict__ arp,
                             const int* __restrict__ aci, const double* __restrict__ av,
                             const int* __restrict__ brp, const int* __restrict__ bci,
                             const double* __restrict__ bv, const int* __restrict__ crp,
                             int* __restrict__ cout_, double* __restrict__ cval,
                             int* __restrict__ err)
    {
        const int r = blockIdx.x;
        if(r >= nrows)
            return;
        const long long o = toff[r];
        const int       H = (int)(toff[r + 1] - o);
        spgemm_row<NUMERIC, true>(rows[r], tkeys + o, NUMERIC ? tvals + o : nullptr, H,
                            threadIdx.x & 63, arp, aci, av, brp, bci, bv, crp, cout_, cval, err);
    }

    // LDS tables: up to 4096 slots with values (numeric pass, 48 KB), up to
    // 32768 keys alone (symbolic pass, 128 KB of the 160 KB of LDS).
    constexpr int kLdsMaxH    = 4096;
    constexpr int kLdsMaxHSym = 32768;

    // Slots of a row's table: the smallest power of two >= 64 holding twice
    // the row's bound on distinct columns (its product or nonzero count,
    // capped by the column count of C).
    __device__ __forceinline__ long long table_slots(int prod, int ncols)
    {
        const long long cap  = ncols > 1 ? ncols : 1;
        const long long need = 2LL * (prod < cap ? prod : cap);
        long long       H    = 64;
        while(H < need)
            H <<= 1;
        return H;
    }

    // Rows of C grouped by hash-table size.
    struct spgemm_plan
    {
        std::vector<int>  bin_start; // per LDS bin (H = 64 << b), into d_rows
        int               n_lds    = 0;
        int               n_global = 0;
        std::vector<long long> toff; // global table offsets (n_global + 1)
        int*              d_rows   = nullptr; // LDS-bin rows, then global rows
        long long*        d_toff   = nullptr;
        int*              d_keys   = nullptr;
        double*           d_vals   = nullptr;

        int release(hipStream_t s)
        {
            if(d_rows)
                LAL_HIP(hipFreeAsync(d_rows, s));
            if(d_toff)
                LAL_HIP(hipFreeAsync(d_toff, s));
            if(d_keys)
                LAL_HIP(hipFreeAsync(d_keys, s));
            if(d_vals)
                LAL_HIP(hipFreeAsync(d_vals, s));
            d_rows = nullptr;
            d_toff = nullptr;
            d_keys = nullptr;
            d_vals = nullptr;
            return LAL_OK;
        }
    };

    constexpr int kBins = 10; // 64 .. 32768 (above 4096 only for the symbolic pass)

    // Bin of row i of C, or -1 for a row the pass skips: bins 0 .. kBins-1
    // hold LDS tables of 64 << b slots, bin kBins tables in global memory
    // (*slots gets the row's table size).  The symbolic pass sizes a row by
    // its product count (`prod`) and skips rows without products; the
    // numeric pass sizes every row by its nonzero count from the symbolic
    // pass (`crp`), empty rows included, so that a row pointer which
    // undercounts them is detected.
    __device__ __forceinline__ int row_bin(int i, const int* __restrict__ prod,
                                           const int* __restrict__ crp, int ncols, int max_h,
                                           long long* slots)
    {
        const int p = crp ? crp[i + 1] - crp[i] : prod[i];
        if(!crp && p <= 0)
            return -1;
        const long long H = table_slots(p > 0 ? p : 0, ncols);
        *slots            = H;
        int b             = 0;
        while((64LL << b) < H && b < kBins)
            b++;
        return H > max_h ? kBins : b;
    }

    // Rows per bin, gathered per workgroup in LDS.
    __global__ void __launch_bounds__(256)
        bin_count_kernel(int m, const int* __restrict__ prod, const int* __restrict__ crp,
                         int ncols, int max_h, int* __restrict__ hist)
    {
        __shared__ int s_n[kBins + 1];
        if(threadIdx.x <= kBins)
            s_n[threadIdx.x] = 0;
        __syncthreads();
        const int i = blockIdx.x * 256 + threadIdx.x;
        long long H = 0;
        const int b = i < m ? row_bin(i, prod, crp, ncols, max_h, &H) : -1;
        if(b >= 0)
            atomicAdd(&s_n[b], 1);
        __syncthreads();
        if(threadIdx.x <= kBins && s_n[threadIdx.x])
            atomicAdd(&hist[threadIdx.x], s_n[threadIdx.x]);
    }

    // Rows into their bins: bin b follows the rows of bins 0 .. b-1, and
    // within a bin the workgroups take their places in order of arrival (each
    // row of C is computed on its own, so the order changes no result).  A
    // row of the global bin also records its table size in gslots.
    __global__ void __launch_bounds__(256)
        bin_scatter_kernel(int m, const int* __restrict__ prod, const int* __restrict__ crp,
                           int ncols, int max_h, const int* __restrict__ hist,
                           int* __restrict__ cursor, int* __restrict__ rows,
                           long long* __restrict__ gslots)
    {
        __shared__ int s_n[kBins + 1], s_base[kBins + 1];
        if(threadIdx.x <= kBins)
            s_n[threadIdx.x] = 0;
        __syncthreads();
        const int i     = blockIdx.x * 256 + threadIdx.x;
        long long H     = 0;
        const int b     = i < m ? row_bin(i, prod, crp, ncols, max_h, &H) : -1;
        const int local = b >= 0 ? atomicAdd(&s_n[b], 1) : 0;
        __syncthreads();
        if(threadIdx.x <= kBins && s_n[threadIdx.x])
        {
            int start = 0;
            for(int q = 0; q < (int)threadIdx.x; q++)
                start += hist[q];
            s_base[threadIdx.x] = start + atomicAdd(&cursor[threadIdx.x], s_n[threadIdx.x]);
        }
        __syncthreads();
        if(b >= 0)
        {
            const int pos = s_base[b] + local;
            rows[pos]     = i;
            if(b == kBins)
            {
                int n_lds = 0;
                for(int q = 0; q < kBins; q++)
                    n_lds += hist[q];
                gslots[pos - n_lds] = H;
            }
        }
    }

    // Rows of C binned by table size on the device; the host reads back the
    // eleven bin counts (and, when some rows need global tables, their table
    // sizes) to size the launches and the global tables.
    int make_plan(lal_context* c, int m, int n, const int* arp, const int* aci, const int* brp,
                  const int* crp, bool numeric, spgemm_plan& P, long nnz_a = 0)
    {
        constexpr int nb = kBins + 1;
        static_assert(nb <= 16, "the bin counts are read back through the context's flags");
        P.bin_start.assign(nb, 0);
        P.n_lds    = 0;
        P.n_global = 0;
        if(m <= 0)
            return LAL_OK;
        const int  max_h  = numeric ? kLdsMaxH : kLdsMaxHSym;
        const int  grid   = lal::grid_for(m, 256);
        int*       d_prod = nullptr;
        int*       d_bins = nullptr; // counts, then cursors
        long long* d_gsl  = nullptr;
        if(!numeric)
        {
            LAL_HIP(hipMallocAsync((void**)&d_prod, sizeof(int) * m, c->stream));
            // rows of A averaging 16 or more entries: 16 lanes per row
            if(nnz_a >= 16L * m)
                products_group_kernel<16><<<(int)(((long)m * 16 + 255) / 256), 256, 0,
                                            c->stream>>>(m, arp, aci, brp, d_prod);
            else
                products_kernel<<<grid, 256, 0, c->stream>>>(m, arp, aci, brp, d_prod);
            LAL_LAUNCHED();
        }
        LAL_HIP(hipMallocAsync((void**)&d_bins, sizeof(int) * 2 * nb, c->stream));
        LAL_HIP(hipMemsetAsync(d_bins, 0, sizeof(int) * 2 * nb, c->stream));
        bin_count_kernel<<<grid, 256, 0, c->stream>>>(m, d_prod, numeric ? crp : nullptr, n,
                                                      max_h, d_bins);
        LAL_LAUNCHED();
        LAL_HIP(hipMemcpyAsync(c->h_flag, d_bins, sizeof(int) * nb, hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        long long total = 0;
        for(int b = 0; b < kBins; b++)
        {
            P.bin_start[b] = (int)total;
            total += c->h_flag[b];
        }
        P.bin_start[kBins] = (int)total;
        P.n_lds            = (int)total;
        P.n_global         = c->h_flag[kBins];
        total += P.n_global;
        if(total)
        {
            LAL_HIP(hipMallocAsync((void**)&P.d_rows, sizeof(int) * total, c->stream));
            if(P.n_global)
                LAL_HIP(hipMallocAsync((void**)&d_gsl, sizeof(long long) * P.n_global,
                                       c->stream));
            bin_scatter_kernel<<<grid, 256, 0, c->stream>>>(m, d_prod, numeric ? crp : nullptr, n,
                                                            max_h, d_bins, d_bins + nb, P.d_rows,
                                                            d_gsl);
            LAL_LAUNCHED();
        }
        if(P.n_global)
        {
            P.toff.assign(P.n_global + 1, 0);
            LAL_HIP(hipMemcpyAsync(P.toff.data() + 1, d_gsl, sizeof(long long) * P.n_global,
                                   hipMemcpyDeviceToHost, c->stream));
            LAL_HIP(hipStreamSynchronize(c->stream));
            for(int r = 0; r < P.n_global; r++)
                P.toff[r + 1] += P.toff[r];
            const long long slots = P.toff[P.n_global];
            LAL_HIP(hipMallocAsync((void**)&P.d_toff, sizeof(long long) * P.toff.size(),
                                   c->stream));
            LAL_HIP(hipMemcpyAsync(P.d_toff, P.toff.data(), sizeof(long long) * P.toff.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_HIP(hipMallocAsync((void**)&P.d_keys, sizeof(int) * slots, c->stream));
            if(numeric)
                LAL_HIP(hipMallocAsync((void**)&P.d_vals, sizeof(double) * slots, c->stream));
            // P.toff must outlive the copy to the device
            LAL_HIP(hipStreamSynchronize(c->stream));
        }
        if(d_prod)
            LAL_HIP(hipFreeAsync(d_prod, c->stream));
        if(d_gsl)
            LAL_HIP(hipFreeAsync(d_gsl, c->stream));
        LAL_HIP(hipFreeAsync(d_bins, c->stream));
        return LAL_OK;
    }

    template <bool NUMERIC>
    int run_plan(lal_context* c, const spgemm_plan& P, const int* arp, const int* aci,
                 const double* av, const int* brp, const int* bci, const double* bv,
                 const int* crp, int* cout_, double* cval)
    {
        for(int b = 0; b < kBins; b++)
        {
            int n0 = P.bin_start[b], nr = P.bin_start[b + 1] - n0;
            if(nr == 0)
                continue;
            const int* rows = P.d_rows + n0;
#define LAL_SPGEMM_GROUP(HH, G, NT, NUM)                                                     \
    spgemm_group_kernel<HH, G, NT, NUM><<<(nr + NT / G - 1) / (NT / G), NT, 0, c->stream>>>(  \
        nr, rows, arp, aci, av, brp, bci, bv, crp, cout_, cval, c->d_flag)
            switch(b)
            {
            case 0: LAL_SPGEMM_GROUP(64, 8, 256, NUMERIC); break;
            case 1: LAL_SPGEMM_GROUP(128, 16, 256, NUMERIC); break;
            case 2: LAL_SPGEMM_GROUP(256, 32, 256, NUMERIC); break;
            // 512 and 1024 slots: a wavefront per row for the numeric pass
            // (24 / 48 KB for four rows), 32 lanes for the symbolic pass's
            // key-only tables (sized by products; 8 rows hold 16 / 32 KB)
            case 3:
                if constexpr(NUMERIC)
                    LAL_SPGEMM_GROUP(512, 64, 256, true);
                else
                    LAL_SPGEMM_GROUP(512, 32, 256, false);
                break;
            case 4:
                if constexpr(NUMERIC)
                    LAL_SPGEMM_GROUP(1024, 64, 256, true);
                else
                    LAL_SPGEMM_GROUP(1024, 32, 256, false);
                break;
            // one row (a wavefront) per workgroup from 2048 slots on
            case 5: LAL_SPGEMM_GROUP(2048, 64, 64, NUMERIC); break;
            case 6: LAL_SPGEMM_GROUP(4096, 64, 64, NUMERIC); break;
            // key-only tables of the symbolic pass (the numeric pass never bins here)
            case 7: LAL_SPGEMM_GROUP(8192, 64, 64, false); break;
            case 8: LAL_SPGEMM_GROUP(16384, 64, 64, false); break;
            case 9: LAL_SPGEMM_GROUP(32768, 64, 64, false); break;
            }
#undef LAL_SPGEMM_GROUP
            LAL_LAUNCHED();
        }
        if(P.n_global)
        {
            spgemm_global_kernel<NUMERIC><<<P.n_global, 64, 0, c->stream>>>(
                P.n_global, P.d_rows + P.n_lds, P.d_toff, P.d_keys, P.d_vals, arp, aci, av, brp,
                bci, bv, crp, cout_, cval, c->d_flag);
            LAL_LAUNCHED();
        }
        return LAL_OK;
    }
}

namespace lal
{
    template <typename T>
    int transpose_t(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
                    const T* v, int* trp, int* tci, T* tv)
    {
        LAL_HIP(hipMemsetAsync(trp, 0, sizeof(int) * (n + 1), c->stream));
        if(nnz > 0)
        {
            col_count_kernel<<<grid_for(nnz, 256), 256, 0, c->stream>>>(nnz, ci, trp);
            LAL_LAUNCHED();
        }
        LAL_TRY(scan_i32_dev(c, n + 1, trp));
        if(nnz == 0)
            return LAL_OK;
        int* cursor = nullptr;
        LAL_HIP(hipMallocAsync((void**)&cursor, sizeof(int) * std::max(n, 1), c->stream));
        LAL_HIP(hipMemcpyAsync(cursor, trp, sizeof(int) * n, hipMemcpyDeviceToDevice, c->stream));
        transpose_scatter_kernel<T><<<grid_for(m, 256), 256, 0, c->stream>>>(m, rp, ci, v,
                                                                             cursor, tci, tv);
        LAL_LAUNCHED();
        row_sort_wave_kernel<T><<<(n + 3) / 4, 256, 0, c->stream>>>(n, trp, tci, tv);
        LAL_LAUNCHED();
        LAL_HIP(hipFreeAsync(cursor, c->stream));
        return LAL_OK;
    }

    int transpose_f64(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
                      const double* v, int* trp, int* tci, double* tv)
    {
        return transpose_t<double>(c, m, n, nnz, rp, ci, v, trp, tci, tv);
    }

    int csrgemm_nnz(lal_context* c, int m, int k, int n, const int* arp, const int* aci,
                    const int* brp, const int* bci, int* crp, int* nnz_c, long nnz_a)
    {
        (void)k;
        LAL_HIP(hipMemsetAsync(crp, 0, sizeof(int) * (m + 1), c->stream));
        spgemm_plan P;
        LAL_TRY(make_plan(c, m, n, arp, aci, brp, nullptr, false, P, nnz_a));
        int st = run_plan<false>(c, P, arp, aci, nullptr, brp, bci, nullptr, nullptr, crp,
                                 nullptr);
        P.release(c->stream);
        LAL_TRY(st);
        LAL_TRY(scan_i32_dev(c, m + 1, crp));
        LAL_HIP(hipMemcpyAsync(c->h_flag, crp + m, sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        *nnz_c = c->h_flag[0];
        return LAL_OK;
    }

    int csrgemm_solve_f64(lal_context* c, int m, int k, int n, const int* arp, const int* aci,
                          const double* av, const int* brp, const int* bci, const double* bv,
                          const int* crp, int* cci, double* cv)
    {
        (void)k;
        spgemm_plan P;
        LAL_TRY(make_plan(c, m, n, arp, aci, brp, crp, true, P));
        LAL_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        int st = run_plan<true>(c, P, arp, aci, av, brp, bci, bv, crp, cci, cv);
        P.release(c->stream);
        LAL_TRY(st);
        LAL_HIP(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        if(c->h_flag[0])
        {
            set_last_error("csrgemm_solve: row_ptr of C does not match the product's pattern");
            return LAL_ERR_INVALID;
        }
        return LAL_OK;
    }
}

namespace
{
    // flag = 1 when a row of B holds a column index twice.  Strictly
    // ascending rows (the usual case) take one pass; other rows compare every
    // pair.
    __global__ void __launch_bounds__(256)
        dup_cols_kernel(int k, const int* __restrict__ brp, const int* __restrict__ bci,
                        int* __restrict__ flag)
    {
        const int row = blockIdx.x * 256 + threadIdx.x;
        if(row >= k)
            return;
        const int rs = brp[row], re = brp[row + 1];
        bool      ascending = true;
        for(int p = rs + 1; p < re && ascending; p++)
            ascending = bci[p] > bci[p - 1];
        if(ascending)
            return;
        for(int p = rs + 1; p < re; p++)
            for(int q = rs; q < p; q++)
                if(bci[q] == bci[p])
                {
                    atomicOr(flag, 1);
                    return;
                }
    }

    // Validates the operands of the C ABI products: column indices of A in
    // [0, k) and of B in [0, n), and no column index repeated within a row of
    // B (the hash pass folds each product into one slot per column, so two
    // lanes meeting the same column of one B row would race; such a matrix is
    // refused rather than summed).
    int check_product_operands(lal_context* c, int k, int n, int nnz_A, const int* aci,
                               int nnz_B, const int* brp, const int* bci)
    {
        LAL_TRY(lal::check_columns(c, nnz_A, k, aci));
        LAL_TRY(lal::check_columns(c, nnz_B, n, bci));
        if(k == 0 || nnz_B == 0)
            return LAL_OK;
        LAL_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int), c->stream));
        dup_cols_kernel<<<lal::grid_for(k, 256), 256, 0, c->stream>>>(k, brp, bci, c->d_flag);
        LAL_LAUNCHED();
        LAL_HIP(hipMemcpyAsync(c->h_flag, c->d_flag, sizeof(int), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        if(c->h_flag[0])
        {
            lal::set_last_error("csrgemm: a row of B repeats a column index");
            return LAL_ERR_UNSUPPORTED;
        }
        return LAL_OK;
    }
}

extern "C" int lal_transpose_f64(lal_handle c, int m, int n, int nnz, const int* rp,
                                 const int* ci, const double* v, int* trp, int* tci, double* tv)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && n >= 0 && nnz >= 0 && trp != nullptr);
    LAL_REQUIRE(nnz == 0 || (rp && ci && v && tci && tv));
    return lal::transpose_f64(c, m, n, nnz, rp, ci, v, trp, tci, tv);
}

extern "C" int lal_transpose_f32(lal_handle c, int m, int n, int nnz, const int* rp,
                                 const int* ci, const float* v, int* trp, int* tci, float* tv)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && n >= 0 && nnz >= 0 && trp != nullptr);
    LAL_REQUIRE(nnz == 0 || (rp && ci && v && tci && tv));
    return lal::transpose_t<float>(c, m, n, nnz, rp, ci, v, trp, tci, tv);
}

extern "C" int lal_csrgemm_nnz(lal_handle c, int m, int k, int n, int nnz_A, const int* arp,
                               const int* aci, int nnz_B, const int* brp, const int* bci,
                               int* crp, int* nnz_C)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && k >= 0 && n >= 0 && nnz_A >= 0 && nnz_B >= 0);
    LAL_REQUIRE(crp != nullptr && nnz_C != nullptr && arp != nullptr && brp != nullptr);
    LAL_REQUIRE((nnz_A == 0 || aci) && (nnz_B == 0 || bci));
    LAL_TRY(check_product_operands(c, k, n, nnz_A, aci, nnz_B, brp, bci));
    return lal::csrgemm_nnz(c, m, k, n, arp, aci, brp, bci, crp, nnz_C, nnz_A);
}

extern "C" int lal_csrgemm_solve_f64(lal_handle c, int m, int k, int n, int nnz_A,
                                     const int* arp, const int* aci, const double* av,
                                     int nnz_B, const int* brp, const int* bci,
                                     const double* bv, const int* crp, int* cci, double* cv)
{
    LAL_REQUIRE(c != nullptr && m >= 0 && k >= 0 && n >= 0 && nnz_A >= 0 && nnz_B >= 0);
    LAL_REQUIRE(crp != nullptr && arp != nullptr && brp != nullptr);
    LAL_REQUIRE((nnz_A == 0 || (aci && av)) && (nnz_B == 0 || (bci && bv)));
    LAL_TRY(check_product_operands(c, k, n, nnz_A, aci, nnz_B, brp, bci));
    return lal::csrgemm_solve_f64(c, m, k, n, arp, aci, av, brp, bci, bv, crp, cci, cv);
}
