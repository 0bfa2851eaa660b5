// Sparse products of the SA-AMG setup: transpose (R = P^T) and the two-pass
// hash SpGEMM used for the Galerkin product A_c = R (A P).
//
// Transpose: count, scan, scatter, then each row of the result is sorted by
// column, which gives the stable order of backend/host/host_csr2csc.cpp:30-75.
//
// SpGEMM: one wavefront per row of C.  Rows are binned by their product count
// into LDS hash tables of H = 64 ... 4096 slots (open addressing, linear
// probing, slots claimed with atomicCAS); rows beyond that get a table in
// global memory.  The numeric pass walks the A row sequentially and lets the
// lanes take the B row in parallel; the columns of one B row are distinct, so
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

    template <int H, int WPB, bool NUMERIC>
    __global__ void __launch_bounds__(WPB * 64)
        spgemm_lds_kernel(int nrows, const int* __restrict__ rows, const int* __restrict__ arp,
                          const int* __restrict__ aci, const double* __restrict__ av,
                          const int* __restrict__ brp, const int* __restrict__ bci,
                          const double* __restrict__ bv, const int* __restrict__ crp,
                          int* __restrict__ cout_, double* __restrict__ cval,
                          int* __restrict__ err)
    {
        __shared__ int    s_keys[WPB][H];
        __shared__ double s_vals[NUMERIC ? WPB : 1][NUMERIC ? H : 1];
        const int         wid  = threadIdx.x >> 6;
        const int         lane = threadIdx.x & 63;
        const int         r    = blockIdx.x * WPB + wid;
        if(r >= nrows)
            return;
        spgemm_row<NUMERIC, false>(rows[r], s_keys[wid], NUMERIC ? s_vals[wid] : nullptr, H, lane, arp,
                            aci, av, brp, bci, bv, crp, cout_, cval, err);
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

    int table_size(int prod, int ncols)
    {
        long long need = 2LL * std::min(prod, std::max(ncols, 1));
        long long H    = 64;
        while(H < need)
            H <<= 1;
        return (int)H;
    }

    // Rows of C grouped by hash-table size.
    struct spgemm_plan
    {
        std::vector<int>  bin_start; // per LDS bin (H = 64 << b), into rows
        std::vector<int>  rows;      // LDS-bin rows followed by global rows
        int               n_lds    = 0;
        int               n_global = 0;
        std::vector<long long> toff; // global table offsets (n_global + 1)
        int*              d_rows   = nullptr;
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

    // Rows of C binned by table size: the symbolic pass sizes a row's table by
    // its product count (an upper bound of its distinct columns), the numeric
    // pass by its exact nonzero count from the symbolic pass (crp), which keeps
    // almost every row in LDS.
    int make_plan(lal_context* c, int m, int n, const int* arp, const int* aci, const int* brp,
                  const int* crp, bool numeric, spgemm_plan& P)
    {
        std::vector<int> prod(m);
        if(numeric)
        {
            std::vector<int> h(m + 1);
            LAL_HIP(hipMemcpyAsync(h.data(), crp, sizeof(int) * (m + 1), hipMemcpyDeviceToHost,
                                   c->stream));
            LAL_HIP(hipStreamSynchronize(c->stream));
            for(int i = 0; i < m; i++)
                prod[i] = h[i + 1] - h[i];
        }
        else
        {
            int* d_prod = nullptr;
            LAL_HIP(hipMallocAsync((void**)&d_prod, sizeof(int) * std::max(m, 1), c->stream));
            products_kernel<<<lal::grid_for(m, 256), 256, 0, c->stream>>>(m, arp, aci, brp,
                                                                          d_prod);
            LAL_LAUNCHED();
            LAL_HIP(hipMemcpyAsync(prod.data(), d_prod, sizeof(int) * m,
                                   hipMemcpyDeviceToHost, c->stream));
            LAL_HIP(hipFreeAsync(d_prod, c->stream));
            LAL_HIP(hipStreamSynchronize(c->stream));
        }

        const int                     max_h = numeric ? kLdsMaxH : kLdsMaxHSym;
        std::vector<std::vector<int>> bins(kBins + 1);
        for(int i = 0; i < m; i++)
        {
            // the numeric pass also visits rows the row pointer says are empty,
            // so that a row pointer which undercounts them is detected
            if(prod[i] <= 0 && !numeric)
                continue;
            int H = table_size(std::max(prod[i], 0), n);
            int b = 0;
            while((64 << b) < H && b < kBins)
                b++;
            bins[H > max_h ? kBins : b].push_back(i);
        }
        P.bin_start.assign(kBins + 1, 0);
        P.rows.clear();
        for(int b = 0; b < kBins; b++)
        {
            P.bin_start[b] = (int)P.rows.size();
            P.rows.insert(P.rows.end(), bins[b].begin(), bins[b].end());
        }
        P.bin_start[kBins] = (int)P.rows.size();
        P.n_lds            = (int)P.rows.size();
        P.rows.insert(P.rows.end(), bins[kBins].begin(), bins[kBins].end());
        P.n_global = (int)bins[kBins].size();
        P.toff.assign(P.n_global + 1, 0);
        for(int r = 0; r < P.n_global; r++)
            P.toff[r + 1] = P.toff[r] + table_size(prod[bins[kBins][r]], n);

        if(!P.rows.empty())
        {
            LAL_HIP(hipMallocAsync((void**)&P.d_rows, sizeof(int) * P.rows.size(), c->stream));
            LAL_HIP(hipMemcpyAsync(P.d_rows, P.rows.data(), sizeof(int) * P.rows.size(),
                                   hipMemcpyHostToDevice, c->stream));
        }
        if(P.n_global)
        {
            long long slots = P.toff[P.n_global];
            LAL_HIP(hipMallocAsync((void**)&P.d_toff, sizeof(long long) * P.toff.size(),
                                   c->stream));
            LAL_HIP(hipMemcpyAsync(P.d_toff, P.toff.data(), sizeof(long long) * P.toff.size(),
                                   hipMemcpyHostToDevice, c->stream));
            LAL_HIP(hipMallocAsync((void**)&P.d_keys, sizeof(int) * slots, c->stream));
            if(numeric)
                LAL_HIP(hipMallocAsync((void**)&P.d_vals, sizeof(double) * slots, c->stream));
        }
        // the host vectors above must outlive the async copies
        LAL_HIP(hipStreamSynchronize(c->stream));
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
#define LAL_SPGEMM_BIN(HH, WPB, NUM)                                                         \
    spgemm_lds_kernel<HH, WPB, NUM><<<(nr + WPB - 1) / WPB, WPB * 64, 0, c->stream>>>(       \
        nr, rows, arp, aci, av, brp, bci, bv, crp, cout_, cval, c->d_flag)
            switch(b)
            {
            case 0: LAL_SPGEMM_BIN(64, 4, NUMERIC); break;
            case 1: LAL_SPGEMM_BIN(128, 4, NUMERIC); break;
            case 2: LAL_SPGEMM_BIN(256, 4, NUMERIC); break;
            case 3: LAL_SPGEMM_BIN(512, 4, NUMERIC); break;
            case 4: LAL_SPGEMM_BIN(1024, 4, NUMERIC); break;
            case 5: LAL_SPGEMM_BIN(2048, 1, NUMERIC); break;
            case 6: LAL_SPGEMM_BIN(4096, 1, NUMERIC); break;
            // key-only tables of the symbolic pass (the numeric pass never bins here)
            case 7: LAL_SPGEMM_BIN(8192, 1, false); break;
            case 8: LAL_SPGEMM_BIN(16384, 1, false); break;
            case 9: LAL_SPGEMM_BIN(32768, 1, false); break;
            }
#undef LAL_SPGEMM_BIN
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
                    const int* brp, const int* bci, int* crp, int* nnz_c)
    {
        (void)k;
        LAL_HIP(hipMemsetAsync(crp, 0, sizeof(int) * (m + 1), c->stream));
        spgemm_plan P;
        LAL_TRY(make_plan(c, m, n, arp, aci, brp, nullptr, false, P));
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
    return lal::csrgemm_nnz(c, m, k, n, arp, aci, brp, bci, crp, nnz_C);
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
