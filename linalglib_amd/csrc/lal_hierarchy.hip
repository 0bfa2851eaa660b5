// SA-AMG hierarchy on the device, the V-cycle preconditioner, SA-AMG-PCG
// and mixed-precision iterative refinement.
//
//   setup      iterative_solvers/amg/saamg.cpp:210-282 (level loop: strength,
//              PMIS, smoothed P, R = P^T, A_c = R (A P); eps halves per level;
//              stop when P has no columns or at min(max_levels, 20) levels)
//   V-cycle    iterative_solvers/amg/amg.cpp:144-267 with apply_smoother(jacobi)
//   PCG        iterative_solvers/krylov/cg.cpp:136-223, convergence test of
//              iterative_solvers/iter_control.cpp:38-91
//
// Everything stays on the device.  Solver scalars (gamma, alpha, beta) are
// formed on the device from per-block partials, so a PCG iteration hands one
// 8-byte residual norm to the host for the convergence test and nothing else.
#include "lal_hier.h"
#include "lal_rowvec.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace lal
{
    int sa_strength(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                    const double* v, double eps, int* conn);
    int pmis_aggregate(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                       const int* conn, int64_t* agg);
    int sa_prolongator_nnz(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                           const int* conn, const int64_t* agg, int* prp, int* nnz_P, int* n_c);
    int sa_prolongator_fill(lal_context* c, int m, int nnz, const int* rp, const int* ci,
                            const double* v, const int* conn, const int64_t* agg, double relax,
                            const int* prp, int* pci, double* pv);
    int ua_prolongator_nnz(lal_context* c, int m, const int64_t* agg, int* prp, int* nnz_P,
                           int* n_c);
    int ua_prolongator_fill(lal_context* c, int m, const int64_t* agg, const int* prp, int* pci,
                            double* pv);
    int transpose_f64(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
                      const double* v, int* trp, int* tci, double* tv);
    // nnz_a: A's nonzeros, which pick the lanes per row of the product count
    int csrgemm_nnz(lal_context* c, int m, int k, int n, const int* arp, const int* aci,
                    const int* brp, const int* bci, int* crp, int* nnz_c, long nnz_a = 0);
    int csrgemm_solve_f64(lal_context* c, int m, int k, int n, const int* arp, const int* aci,
                          const double* av, const int* brp, const int* bci, const double* bv,
                          const int* crp, int* cci, double* cv);
}

using lal::alloc_n;
using lal::converges;
using lal::post_scalar;
using lal::wait_scalar;
using lal::csr_dev;
using lal::level_t;
using lal::work;

namespace
{
    template <typename T>
    const T* vals(const csr_dev& A);
    template <>
    const double* vals<double>(const csr_dev& A)
    {
        return A.v;
    }
    template <>
    const float* vals<float>(const csr_dev& A)
    {
        return A.v32;
    }

    template <typename T>
    const lal::tail_desc& tail_of(const lal_hierarchy_t* h);
    template <>
    const lal::tail_desc& tail_of<double>(const lal_hierarchy_t* h)
    {
        return h->tail64;
    }
    template <>
    const lal::tail_desc& tail_of<float>(const lal_hierarchy_t* h)
    {
        return h->tail32;
    }
    template <typename T>
    const int* tail_image(const lal_hierarchy_t* h);
    template <>
    const int* tail_image<double>(const lal_hierarchy_t* h)
    {
        return h->tail_img64;
    }
    template <>
    const int* tail_image<float>(const lal_hierarchy_t* h)
    {
        return h->tail_img32;
    }

    // r = b - A x in the level's precision
    template <typename T>
    int residual_t(lal_context* c, const csr_dev& A, const T* Av, const T* x, const T* b, T* r)
    {
        if(sizeof(T) == 8)
            return lal::residual_f64(c, A.d, A.m, A.n, A.nnz, A.rp, A.ci, (const double*)Av,
                                     (const double*)x, (const double*)b, (double*)r);
        return lal::residual_f32(c, A.d, A.m, A.n, A.nnz, A.rp, A.ci, (const float*)Av,
                                 (const float*)x, (const float*)b, (float*)r);
    }

    // Dense inverse of the coarsest operator through its Cholesky factor
    // (FP64, on the host; the operator is small).  Returns false if a pivot
    // is not positive.
    bool dense_inverse(int n, const std::vector<int>& rp, const std::vector<int>& ci,
                       const std::vector<double>& v, std::vector<double>& inv)
    {
        std::vector<double> Lf((size_t)n * n, 0.0);
        for(int i = 0; i < n; i++)
            for(int j = rp[i]; j < rp[i + 1]; j++)
                Lf[(size_t)i * n + ci[j]] = v[j];
        for(int j = 0; j < n; j++)
        {
            double s = Lf[(size_t)j * n + j];
            for(int k = 0; k < j; k++)
                s -= Lf[(size_t)j * n + k] * Lf[(size_t)j * n + k];
            if(!(s > 0.0))
                return false;
            double d             = std::sqrt(s);
            Lf[(size_t)j * n + j] = d;
            for(int i = j + 1; i < n; i++)
            {
                double t = Lf[(size_t)i * n + j];
                for(int k = 0; k < j; k++)
                    t -= Lf[(size_t)i * n + k] * Lf[(size_t)j * n + k];
                Lf[(size_t)i * n + j] = t / d;
            }
        }
        inv.assign((size_t)n * n, 0.0);
        std::vector<double> col(n);
        for(int e = 0; e < n; e++)
        {
            for(int i = 0; i < n; i++)
            {
                double t = (i == e) ? 1.0 : 0.0;
                for(int k = 0; k < i; k++)
                    t -= Lf[(size_t)i * n + k] * col[k];
                col[i] = t / Lf[(size_t)i * n + i];
            }
            for(int i = n - 1; i >= 0; i--)
            {
                double t = col[i];
                for(int k = i + 1; k < n; k++)
                    t -= Lf[(size_t)k * n + i] * col[k];
                col[i] = t / Lf[(size_t)i * n + i];
            }
            for(int i = 0; i < n; i++)
                inv[(size_t)i * n + e] = col[i];
        }
        return true;
    }

    // Layout of the coarse tail's LDS image for levels first..levels in
    // precision T (word offsets; values 8-byte aligned for FP64).  With
    // `img` the arrays are also copied into a fresh device image.  Returns
    // false when the levels do not fit the tail's LDS.
    template <typename T>
    bool tail_layout(lal_hierarchy_t* h, int first, lal::tail_desc* d, int** img, int* status)
    {
        constexpr int S = (int)sizeof(T) / 4;
        *status         = LAL_OK;
        const int nl    = h->levels - first + 1;
        if(first < 1 || nl < 2 || nl > lal::TAIL_MAX_LEVELS)
            return false;
        std::memset(d, 0, sizeof(*d));
        d->nl   = nl;
        d->kind = h->coarse_kind == 0 ? 0 : 1;
        long w  = 0;
        struct piece
        {
            long        off;
            const void* src;
            size_t      bytes;
        };
        std::vector<piece> pieces;
        auto put = [&](const void* src, long words, int align) -> int {
            w            = (w + align - 1) / align * align;
            const long o = w;
            if(words > 0)
                pieces.push_back({o, src, (size_t)words * 4});
            w += words;
            return (int)o;
        };
        for(int k = 0; k < nl; k++)
        {
            const int       l  = first + k;
            level_t&        L  = h->L[l];
            work<T>&        W  = h->w<T>(l);
            lal::tail_level& t = d->L[k];
            t.m                = L.A.m;
            t.rp               = put(L.A.rp, L.A.m + 1, 1);
            t.ci               = put(L.A.ci, L.A.nnz, 1);
            t.v                = put(vals<T>(L.A), (long)L.A.nnz * S, S);
            t.dinv             = put(W.dinv, (long)L.A.m * S, S);
            t.prp = t.pci = t.pv = t.rrp = t.rci = t.rv = -1;
            if(l < h->levels)
            {
                t.prp = put(L.P.rp, L.P.m + 1, 1);
                t.pci = put(L.P.ci, L.P.nnz, 1);
                t.pv  = put(vals<T>(L.P), (long)L.P.nnz * S, S);
                t.rrp = put(L.R.rp, L.R.m + 1, 1);
                t.rci = put(L.R.ci, L.R.nnz, 1);
                t.rv  = put(vals<T>(L.R), (long)L.R.nnz * S, S);
            }
        }
        // the coarsest dense inverse joins the image when it fits
        d->inv            = -1;
        const long budget = lal::TAIL_LDS_BYTES / 4;
        long       vec    = 0;
        for(int k = 0; k < nl; k++)
            vec += 4 * ((long)d->L[k].m * S + S);
        const int mc = h->L[h->levels].A.m;
        if(d->kind == 1 && h->coarse_inv)
        {
            const long inv_words = 2L * mc * mc;
            if(((w + 1) / 2 * 2) + inv_words + vec + 8 <= budget)
                d->inv = put(h->coarse_inv, inv_words, 2);
        }
        w            = (w + 3) / 4 * 4;
        d->img_words = (int)w;
        for(int k = 0; k < nl; k++)
        {
            lal::tail_level& t = d->L[k];
            t.x                = put(nullptr, (long)t.m * S, S);
            t.t                = put(nullptr, (long)t.m * S, S);
            t.b                = put(nullptr, (long)t.m * S, S);
            t.r                = put(nullptr, (long)t.m * S, S);
        }
        d->lds_words = (int)w;
        if(w > budget)
            return false;
        if(img)
        {
            *status = alloc_n(img, (size_t)d->img_words);
            if(*status != LAL_OK)
                return false;
            for(const piece& p : pieces)
                if(p.src)
                {
                    hipError_t e = hipMemcpyAsync(*img + p.off, p.src, p.bytes,
                                                  hipMemcpyDeviceToDevice, h->c->stream);
                    if(e != hipSuccess)
                    {
                        *status = lal::hip_status(e, "tail image", __FILE__, __LINE__);
                        return false;
                    }
                }
        }
        return true;
    }

    int to_f32(lal_context* c, const double* x, float** y, size_t n)
    {
        LAL_TRY(alloc_n(y, n));
        if(n)
            LAL_TRY(lal_convert_f64_f32(c, x, *y, n));
        return LAL_OK;
    }

    int finish_setup(lal_hierarchy_t* h)
    {
        lal_context* c     = h->c;
        const int    exact = h->opts.exact;
        for(int l = 0; l <= h->levels; l++)
        {
            level_t& L = h->L[l];
            // level 0 carries the descriptor of the input's validation (the
            // same row pointer, copied); the rest are this setup's products
            if(!L.A.d)
                LAL_TRY(lal::csrmv_analysis_dev(c, L.A.m, L.A.n, L.A.nnz, L.A.rp, L.A.ci, exact,
                                                &L.A.d, true));
            L.A.d->exact = exact;
            // The finest operator is read three times per PCG iteration
            // (residual, post-sweep, the Krylov product), the last two with
            // only vector passes between them: it may stay in the Infinity
            // Cache when it fits (decided per precision at launch).  Keeping
            // a coarser level as well overflows the cache and slows the solve.
            L.A.d->keep = l == 0;
            // the fast path reads the hierarchy's operators with 16-bit
            // columns (offsets from each row block's smallest column) where
            // the blocks allow it: 2 bytes per nonzero less, same results
            if(!exact)
                LAL_TRY(lal::compact_columns(c, L.A.d, L.A.nnz, L.A.ci));
            LAL_TRY(alloc_n(&L.w64.x, L.A.m));
            LAL_TRY(alloc_n(&L.w64.t, L.A.m));
            LAL_TRY(alloc_n(&L.w64.b, L.A.m));
            LAL_TRY(alloc_n(&L.w64.r, L.A.m));
            LAL_TRY(alloc_n(&L.w64.diag, L.A.m));
            LAL_TRY(alloc_n(&L.w64.dinv, L.A.m));
            LAL_TRY(lal::diag_inv<double>(c, L.A.m, L.A.rp, L.A.ci, L.A.v, L.w64.diag,
                                          L.w64.dinv, L.A.nnz));
            if(l < h->levels)
            {
                LAL_TRY(lal::csrmv_analysis_dev(c, L.P.m, L.P.n, L.P.nnz, L.P.rp, L.P.ci, exact,
                                                &L.P.d, true));
                LAL_TRY(lal::csrmv_analysis_dev(c, L.R.m, L.R.n, L.R.nnz, L.R.rp, L.R.ci, exact,
                                                &L.R.d, true));
                if(!exact)
                {
                    LAL_TRY(lal::compact_columns(c, L.P.d, L.P.nnz, L.P.ci));
                    LAL_TRY(lal::compact_columns(c, L.R.d, L.R.nnz, L.R.ci));
                }
            }
            if(h->opts.build_f32)
            {
                LAL_TRY(to_f32(c, L.A.v, &L.A.v32, L.A.nnz));
                if(l < h->levels)
                {
                    LAL_TRY(to_f32(c, L.P.v, &L.P.v32, L.P.nnz));
                    LAL_TRY(to_f32(c, L.R.v, &L.R.v32, L.R.nnz));
                }
                LAL_TRY(alloc_n(&L.w32.x, L.A.m));
                LAL_TRY(alloc_n(&L.w32.t, L.A.m));
                LAL_TRY(alloc_n(&L.w32.b, L.A.m));
                LAL_TRY(alloc_n(&L.w32.r, L.A.m));
                LAL_TRY(alloc_n(&L.w32.diag, L.A.m));
                LAL_TRY(alloc_n(&L.w32.dinv, L.A.m));
                LAL_TRY(lal::diag_inv<float>(c, L.A.m, L.A.rp, L.A.ci, L.A.v32, L.w32.diag,
                                             L.w32.dinv, L.A.nnz));
            }
        }
        // coarsest level
        csr_dev& Ac = h->L[h->levels].A;
        if(Ac.m == 1)
        {
            if(Ac.nnz < 1)
            {
                lal::set_last_error("coarsest 1x1 operator has no entry");
                return LAL_ERR_UNSUPPORTED;
            }
            h->coarse_kind = 0;
        }
        else if(h->opts.coarse_solver == 1 && Ac.m <= 4096)
        {
            std::vector<int>    rp(Ac.m + 1), ci(Ac.nnz);
            std::vector<double> v(Ac.nnz), inv;
            LAL_HIP(hipMemcpyAsync(rp.data(), Ac.rp, sizeof(int) * (Ac.m + 1),
                                   hipMemcpyDeviceToHost, c->stream));
            if(Ac.nnz)
            {
                LAL_HIP(hipMemcpyAsync(ci.data(), Ac.ci, sizeof(int) * Ac.nnz,
                                       hipMemcpyDeviceToHost, c->stream));
                LAL_HIP(hipMemcpyAsync(v.data(), Ac.v, sizeof(double) * Ac.nnz,
                                       hipMemcpyDeviceToHost, c->stream));
            }
            LAL_HIP(hipStreamSynchronize(c->stream));
            if(dense_inverse(Ac.m, rp, ci, v, inv))
            {
                LAL_TRY(alloc_n(&h->coarse_inv, inv.size()));
                LAL_HIP(hipMemcpyAsync(h->coarse_inv, inv.data(), sizeof(double) * inv.size(),
                                       hipMemcpyHostToDevice, c->stream));
                LAL_HIP(hipStreamSynchronize(c->stream));
                h->coarse_kind = 2;
            }
            else
                h->coarse_kind = 1;
        }
        else
            h->coarse_kind = 1;

        // coarse tail: the deepest run of levels (at least two, never level 0)
        // whose FP64 arrays fit the tail's LDS, on the fast path with a direct
        // coarsest solve
        if(!exact && h->coarse_kind != 1 && h->levels >= 2)
        {
            lal::tail_desc probe{};
            int            st    = LAL_OK;
            int            first = -1;
            for(int f = h->levels - 1; f >= 1; f--)
            {
                if(!tail_layout<double>(h, f, &probe, nullptr, &st))
                    break;
                first = f;
            }
            if(first >= 1)
            {
                if(!tail_layout<double>(h, first, &h->tail64, &h->tail_img64, &st))
                    return st != LAL_OK ? st : LAL_ERR_INVALID;
                if(h->opts.build_f32 &&
                   !tail_layout<float>(h, first, &h->tail32, &h->tail_img32, &st))
                    return st != LAL_OK ? st : LAL_ERR_INVALID;
                h->tail_first = first;
                LAL_HIP(hipStreamSynchronize(h->c->stream));
            }
        }

        const int n = h->L[0].A.m;
        LAL_TRY(alloc_n(&h->res, n));
        LAL_TRY(alloc_n(&h->z, n));
        LAL_TRY(alloc_n(&h->p, n));
        LAL_TRY(alloc_n(&h->q, n));
        LAL_TRY(alloc_n(&h->s, lal::LAL_SOLVER_SLOTS));
        LAL_TRY(alloc_n(&h->part, 4 * lal::PCG_BLOCKS));
        LAL_HIP(hipHostMalloc((void**)&h->h_note, 64, hipHostMallocMapped | hipHostMallocCoherent));
        std::memset(h->h_note, 0, 64);
        h->h_seq = reinterpret_cast<unsigned*>(h->h_note + 1);
        LAL_HIP(hipHostGetDevicePointer((void**)&h->d_note, h->h_note, 0));
        h->d_seq = reinterpret_cast<unsigned*>(h->d_note + 1);
        if(h->opts.build_f32)
        {
            LAL_TRY(alloc_n(&h->res32, n));
            LAL_TRY(alloc_n(&h->z32, n));
            LAL_TRY(alloc_n(&h->p32, n));
            LAL_TRY(alloc_n(&h->q32, n));
            LAL_TRY(alloc_n(&h->d32, n));
            h->has32 = true;
        }
        LAL_HIP(hipStreamSynchronize(c->stream));
        return LAL_OK;
    }

    // Level loop of saamg_setup (saamg.cpp:210-282) or, with unsmoothed set,
    // uaamg_setup (uaamg.cpp:125-224): the same strength, PMIS and Galerkin
    // steps around a different prolongator.
    int setup(lal_context* c, int m, int n, int nnz, const int* rp, const int* ci,
              const double* v, const lal_saamg_options& o, lal_hierarchy_t* h,
              bool unsmoothed = false)
    {
        h->c    = c;
        h->opts = o;
        csr_dev& A0 = h->L[0].A;
        A0.m = m;
        A0.n = n;
        A0.nnz = nnz;
        LAL_TRY(alloc_n(&A0.rp, m + 1));
        LAL_TRY(alloc_n(&A0.ci, nnz));
        LAL_TRY(alloc_n(&A0.v, nnz));
        LAL_HIP(hipMemcpyAsync(A0.rp, rp, sizeof(int) * (m + 1), hipMemcpyDeviceToDevice,
                               c->stream));
        if(nnz)
        {
            LAL_HIP(hipMemcpyAsync(A0.ci, ci, sizeof(int) * nnz, hipMemcpyDeviceToDevice,
                                   c->stream));
            LAL_HIP(hipMemcpyAsync(A0.v, v, sizeof(double) * nnz, hipMemcpyDeviceToDevice,
                                   c->stream));
        }
        int    max_level = std::min(o.max_levels, 20);
        double eps       = o.eps;
        int    level     = 0;
        while(level < max_level)
        {
            csr_dev& Af = h->L[level].A;
            csr_dev& P  = h->L[level].P;
            csr_dev& R  = h->L[level].R;
            csr_dev& Ac = h->L[level + 1].A;
            int*     conn = nullptr;
            int64_t* agg  = nullptr;
            LAL_TRY(alloc_n(&conn, Af.nnz));
            LAL_TRY(alloc_n(&agg, Af.m));
            LAL_TRY(lal::sa_strength(c, Af.m, Af.nnz, Af.rp, Af.ci, Af.v, eps, conn));
            LAL_TRY(lal::pmis_aggregate(c, Af.m, Af.nnz, Af.rp, Af.ci, conn, agg));
            LAL_TRY(alloc_n(&P.rp, Af.m + 1));
            int nnzP = 0, nc = 0;
            if(unsmoothed)
                LAL_TRY(lal::ua_prolongator_nnz(c, Af.m, agg, P.rp, &nnzP, &nc));
            else
                LAL_TRY(lal::sa_prolongator_nnz(c, Af.m, Af.nnz, Af.rp, Af.ci, conn, agg, P.rp,
                                                &nnzP, &nc));
            if(nc == 0)
            {
                (void)hipFree(conn);
                (void)hipFree(agg);
                P.free_all();
                break;
            }
            P.m   = Af.m;
            P.n   = nc;
            P.nnz = nnzP;
            LAL_TRY(alloc_n(&P.ci, nnzP));
            LAL_TRY(alloc_n(&P.v, nnzP));
            if(unsmoothed)
                LAL_TRY(lal::ua_prolongator_fill(c, Af.m, agg, P.rp, P.ci, P.v));
            else
                LAL_TRY(lal::sa_prolongator_fill(c, Af.m, Af.nnz, Af.rp, Af.ci, Af.v, conn, agg,
                                                 o.relax, P.rp, P.ci, P.v));
            (void)hipStreamSynchronize(c->stream);
            (void)hipFree(conn);
            (void)hipFree(agg);
            // R = P^T
            R.m   = nc;
            R.n   = Af.m;
            R.nnz = nnzP;
            LAL_TRY(alloc_n(&R.rp, nc + 1));
            LAL_TRY(alloc_n(&R.ci, nnzP));
            LAL_TRY(alloc_n(&R.v, nnzP));
            LAL_TRY(lal::transpose_f64(c, P.m, P.n, P.nnz, P.rp, P.ci, P.v, R.rp, R.ci, R.v));
            // AP = A P
            csr_dev AP;
            AP.m = Af.m;
            AP.n = nc;
            LAL_TRY(alloc_n(&AP.rp, Af.m + 1));
            LAL_TRY(lal::csrgemm_nnz(c, Af.m, Af.n, nc, Af.rp, Af.ci, P.rp, P.ci, AP.rp,
                                     &AP.nnz, Af.nnz));
            LAL_TRY(alloc_n(&AP.ci, AP.nnz));
            LAL_TRY(alloc_n(&AP.v, AP.nnz));
            LAL_TRY(lal::csrgemm_solve_f64(c, Af.m, Af.n, nc, Af.rp, Af.ci, Af.v, P.rp, P.ci,
                                           P.v, AP.rp, AP.ci, AP.v));
            // A_c = R (A P)
            Ac.m = nc;
            Ac.n = nc;
            LAL_TRY(alloc_n(&Ac.rp, nc + 1));
            LAL_TRY(lal::csrgemm_nnz(c, nc, Af.m, nc, R.rp, R.ci, AP.rp, AP.ci, Ac.rp, &Ac.nnz,
                                     R.nnz));
            LAL_TRY(alloc_n(&Ac.ci, Ac.nnz));
            LAL_TRY(alloc_n(&Ac.v, Ac.nnz));
            LAL_TRY(lal::csrgemm_solve_f64(c, nc, Af.m, nc, R.rp, R.ci, R.v, AP.rp, AP.ci, AP.v,
                                           Ac.rp, Ac.ci, Ac.v));
            (void)hipStreamSynchronize(c->stream);
            AP.free_all();
            level++;
            eps *= 0.5;
        }
        h->levels = level;
        return finish_setup(h);
    }

    // One V-cycle at level l for rhs b with the initial guess zero.  The
    // result lands in dst when given, else in a level buffer; *out gets it.
    template <typename T>
    int vcycle(lal_hierarchy_t* h, int l, const T* b, T* dst, int n1, int n2, bool xinit,
               T** out)
    {
        lal_context* c = h->c;
        level_t&     L = h->L[l];
        work<T>&     W = h->w<T>(l);
        const csr_dev& A = L.A;
        const T*       Av = vals<T>(A);
        // Fast path with one pre-sweep: the sweep from zero is x = b/a_ii,
        // written by whoever produced b (the restriction above, or the PCG
        // update for level 0) when xinit is set; this level then only forms
        // r = b - A x.
        const bool fast = !h->opts.exact && n1 == 1;
        if(fast && n2 == 1 && l == h->tail_first && tail_image<T>(h))
        {
            T* res = dst ? dst : W.t;
            LAL_TRY(lal::coarse_tail<T>(c, tail_of<T>(h), tail_image<T>(h), b, h->coarse_inv,
                                        res));
            *out = res;
            return LAL_OK;
        }
        if(l == h->levels)
        {
            T* x = dst ? dst : W.x;
            if(h->coarse_kind == 0)
                LAL_TRY(lal::coarse_scalar<T>(c, Av, b, x));
            else if(h->coarse_kind == 2)
                LAL_TRY(lal::coarse_dense<T>(c, A.m, h->coarse_inv, b, x));
            else
                LAL_TRY(lal::coarse_gs<T>(c, A.m, A.rp, A.ci, Av, b, x));
            *out = x;
            return LAL_OK;
        }
        T* cur   = W.x;
        T* other = W.t;
        if(n1 == 0)
        {
            // x = 0 and r = b - A*0, which is b bit for bit
            LAL_HIP(hipMemsetAsync(cur, 0, sizeof(T) * A.m, c->stream));
            LAL_HIP(hipMemcpyAsync(W.r, b, sizeof(T) * A.m, hipMemcpyDeviceToDevice, c->stream));
        }
        else if(fast)
        {
            if(!xinit)
                LAL_TRY(lal::mul_elem<T>(c, A.m, b, W.dinv, cur));
            LAL_TRY(residual_t<T>(c, A, Av, cur, b, W.r));
        }
        else
        {
            LAL_TRY(lal::jacobi0_residual<T>(c, A.d, A.m, A.rp, A.ci, Av, W.diag, W.dinv, b, cur,
                                             W.r));
            if(n1 > 1)
            {
                for(int k = 1; k < n1; k++)
                {
                    LAL_TRY(lal::jacobi_sweep<T>(c, A.d, A.m, A.rp, A.ci, Av, W.diag, W.dinv,
                                                 cur, b, other));
                    std::swap(cur, other);
                }
                LAL_TRY(residual_t<T>(c, A, Av, cur, b, W.r));
            }
        }
        // restriction: b_{l+1} = R r (and, on the fast path, x_{l+1} = b_{l+1}/a_ii)
        const csr_dev& R  = L.R;
        work<T>&       Wc = h->w<T>(l + 1);
        if(fast)
            LAL_TRY(lal::restrict_init<T>(c, R.d, R.m, R.rp, R.ci, vals<T>(R), W.r, Wc.dinv,
                                          Wc.b, Wc.x));
        else
            LAL_TRY(lal::csrmv(c, R.d, R.m, R.n, R.nnz, 1.0, R.rp, R.ci, vals<T>(R), W.r, 0.0,
                               Wc.b));
        T* ec = nullptr;
        LAL_TRY(vcycle<T>(h, l + 1, Wc.b, nullptr, n1, n2, fast, &ec));
        // prolongation and correction: x += P ec
        const csr_dev& P = L.P;
        LAL_TRY(lal::csrmv(c, P.d, P.m, P.n, P.nnz, 1.0, P.rp, P.ci, vals<T>(P), ec, 1.0, cur));
        for(int k = 0; k < n2; k++)
        {
            T* o = (k == n2 - 1 && dst) ? dst : other;
            LAL_TRY(lal::jacobi_sweep<T>(c, A.d, A.m, A.rp, A.ci, Av, W.diag, W.dinv, cur, b,
                                         o));
            other = cur;
            cur   = o;
        }
        if(dst && cur != dst)
        {
            LAL_HIP(hipMemcpyAsync(dst, cur, sizeof(T) * A.m, hipMemcpyDeviceToDevice,
                                   c->stream));
            cur = dst;
        }
        *out = cur;
        return LAL_OK;
    }

    template <typename T>
    int precond(lal_hierarchy_t* h, const T* r, T* z, int n1, int n2, bool xinit = false)
    {
        T* out = nullptr;
        return vcycle<T>(h, 0, r, z, n1, n2, xinit, &out);
    }

    // precond() for the solvers' own vectors (r, z fixed per hierarchy and
    // precision): with at least one pre-sweep the V-cycle is kernels and
    // device copies only.  On a small finest level (at most GRAPH_MAX_ROWS
    // rows), where every kernel is short and the launches bound the cycle, it
    // is captured as a HIP graph on its second use and replayed after
    // (lal::captured): 2D Poisson 128^2 solves in 2.83 ms against 2.99.  On
    // larger ones the graph's replay costs more than it saves (3D Poisson
    // 128^3: FP64 15.37 against 15.13 ms), so they launch eagerly.
    constexpr int GRAPH_MAX_ROWS = 262144;

    template <typename T>
    int precond_graphed(lal_hierarchy_t* h, const T* r, T* z, int n1, int n2,
                        bool xinit = false)
    {
        if(n1 < 1 || h->L[0].A.m > GRAPH_MAX_ROWS)
            return precond<T>(h, r, z, n1, n2, xinit);
        const int kind = lal::KIND_PRECOND + 2 * (int)xinit + (sizeof(T) == 4 ? 1 : 0);
        return lal::captured(h, z, r, n1, n2, 0, 0, kind,
                             [&] { return precond<T>(h, r, z, n1, n2, xinit); });
    }

    // cg_solver::solve_precond on the device.  T = double uses the FP64
    // levels and workspaces, T = float the FP32 ones.  The Krylov products
    // use `op` when given (a caller's matrix of the finest level's size, FP64
    // only), else the matrix the hierarchy was built from.
    template <typename T>
    int pcg(lal_hierarchy_t* h, const T* b, T* x, T* res, T* z, T* p, T* q, double rel_tol,
            double abs_tol, int max_iter, int n1, int n2, lal_solve_info* info,
            const csr_dev* op = nullptr, int smoother = 0)
    {
        lal_context*   c  = h->c;
        const csr_dev& A  = op ? *op : h->L[0].A;
        const int      n  = A.m;
        const T*       Av = vals<T>(A);
        double*        s  = h->s;
        if(sizeof(T) == 8)
            LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, (const double*)Av,
                                      (const double*)x, (const double*)b, (double*)res));
        else
            LAL_TRY(lal::residual_f32(c, A.d, n, A.n, A.nnz, A.rp, A.ci, (const float*)Av,
                                      (const float*)x, (const float*)b, (float*)res));
        // partial buffers: <q,p>, <z,r> (two generations), max|r|
        double*  part_pq   = h->part;
        double*  part_g[2] = {h->part + lal::PCG_BLOCKS, h->part + 2 * lal::PCG_BLOCKS};
        double*  part_r    = h->part + 3 * lal::PCG_BLOCKS;
        unsigned seq       = 0;
        double   init      = 0.0;
        LAL_TRY(lal::norm_partials<T>(c, n, res, part_r));
        LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM0, &seq));
        LAL_TRY(wait_scalar(h, seq, lal::S_RNORM0, &init));
        // the preconditioner: one V-cycle from zero, with the Jacobi fast path
        // or (smoother != 0, FP64 only) another smoother of the AMG cycles
        auto apply = [&](bool xinit) -> int {
            if(smoother != 0)
                return lal::vcycle_smoothed(h, (const double*)res, (double*)z, n1, n2, smoother);
            return precond_graphed<T>(h, res, z, n1, n2, xinit);
        };
        LAL_TRY(apply(false));
        LAL_HIP(hipMemcpyAsync(p, z, sizeof(T) * n, hipMemcpyDeviceToDevice, c->stream));
        int cur = 0;
        LAL_TRY(lal::dot_partials<T, T>(c, n, z, res, part_g[cur]));
        double rn   = init;
        int    iter = 0;
        bool   conv = false;
        double r_prev = init, rate = 0.5; // last norm, observed reduction per step
        while(!(iter > max_iter))
        {
            LAL_TRY(lal::csrmv(c, A.d, n, A.n, A.nnz, 1.0, A.rp, A.ci, Av, p, 0.0, q));
            LAL_TRY(lal::dot_partials<T, T>(c, n, q, p, part_pq));
            // with one pre-sweep on the fast path, the update also writes the
            // V-cycle's first Jacobi iterate x0 = r / a_ii
            const bool fast = !h->opts.exact && n1 == 1 && h->levels > 0 && smoother == 0;
            work<T>&   W0   = h->w<T>(0);
            LAL_TRY(lal::update_xr<T, T>(c, n, p, q, x, res, part_pq, part_g[cur], part_r,
                                         s + lal::S_ALPHA, W0.dinv, fast ? W0.x : nullptr,
                                         nullptr));
            LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM, &seq));
            // The next preconditioner application depends on r only: it is
            // queued before the host looks at the norm, so the GPU does not
            // wait for the convergence test -- unless the observed rate of
            // reduction predicts that this step converges, when the host
            // decides first so that no V-cycle is spent after the last step.
            const bool near = lal::near_tolerance(r_prev, rate, init, rel_tol, abs_tol);
            if(!near)
                LAL_TRY(apply(fast));
            LAL_TRY(wait_scalar(h, seq, lal::S_RNORM, &rn));
            lal::observe_rate(rn, &r_prev, &rate);
            if(std::isnan(rn))
                break; // a NaN in the residual: not converged
            if(converges(rn, init, rel_tol, abs_tol))
            {
                conv = true;
                break;
            }
            if(near)
                LAL_TRY(apply(fast));
            LAL_TRY(lal::dot_partials<T, T>(c, n, z, res, part_g[cur ^ 1]));
            LAL_TRY(lal::update_p<T, T>(c, n, z, p, part_g[cur ^ 1], part_g[cur]));
            cur ^= 1;
            iter++;
        }
        if(info)
        {
            info->iterations       = iter;
            info->converged        = conv;
            info->initial_res_norm = init;
            info->final_res_norm   = rn;
        }
        return LAL_OK;
    }

    // cg_solver::solve_precond with the Krylov vectors (x, r, p, q), the
    // operator product and every inner product in FP64 and the V-cycle on the
    // FP32 levels.  The residual's FP32 copy and the V-cycle's first Jacobi
    // iterate are written by the FP64 update kernel, and <z, r> / p = z + beta p
    // read the FP32 z directly, so the precision change costs no extra pass.
    // The V-cycle input is r scaled by sigma = 2^-k, k the binary exponent of
    // ||r0||_inf, and rounded to FP32, so its entries stay in FP32's range
    // whatever b's magnitude.  PCG is invariant to a constant scaling of the
    // preconditioner (alpha takes 1/sigma, beta is unchanged), and a power of
    // two scales every iterate exactly, so the iterates are those of the
    // unscaled V-cycle.
    int pcg_mixed(lal_hierarchy_t* h, const double* b, double* x, double rel_tol,
                  double abs_tol, int max_iter, int n1, int n2, lal_solve_info* info)
    {
        lal_context*   c   = h->c;
        const csr_dev& A   = h->L[0].A;
        const int      n   = A.m;
        double*        s   = h->s;
        double*        res = h->res;
        double*        p   = h->p;
        double*        q   = h->q;
        float*         r32 = h->res32;
        float*         z32 = h->z32;
        LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, A.v, x, b, res));
        double*  part_pq   = h->part;
        double*  part_g[2] = {h->part + lal::PCG_BLOCKS, h->part + 2 * lal::PCG_BLOCKS};
        double*  part_r    = h->part + 3 * lal::PCG_BLOCKS;
        unsigned seq       = 0;
        double   init      = 0.0;
        LAL_TRY(lal::norm_partials<double>(c, n, res, part_r));
        LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM0, &seq));
        LAL_TRY(wait_scalar(h, seq, lal::S_RNORM0, &init));
        const double sigma = init > 0.0 && std::isfinite(init) ? std::ldexp(1.0, -std::ilogb(init))
                                                               : 1.0;
        LAL_TRY(lal::scale_cast_f64_f32(c, n, res, sigma, r32));
        LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2));
        LAL_TRY(lal_convert_f32_f64(c, z32, p, (size_t)n));
        int cur = 0;
        LAL_TRY(lal::dot_partials<float, double>(c, n, z32, res, part_g[cur]));
        double       rn   = init;
        int          iter = 0;
        bool         conv = false;
        const bool   fast = !h->opts.exact && n1 == 1 && h->levels > 0;
        work<float>& W0   = h->w<float>(0);
        double       r_prev = init, rate = 0.5;
        while(!(iter > max_iter))
        {
            LAL_TRY(lal::csrmv(c, A.d, n, A.n, A.nnz, 1.0, A.rp, A.ci, A.v, p, 0.0, q));
            LAL_TRY(lal::dot_partials<double, double>(c, n, q, p, part_pq));
            LAL_TRY(lal::update_xr<double, float>(c, n, p, q, x, res, part_pq, part_g[cur],
                                                  part_r, s + lal::S_ALPHA, W0.dinv,
                                                  fast ? W0.x : nullptr, r32, sigma));
            LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM, &seq));
            const bool near = lal::near_tolerance(r_prev, rate, init, rel_tol, abs_tol);
            if(!near)
                LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2, fast));
            LAL_TRY(wait_scalar(h, seq, lal::S_RNORM, &rn));
            lal::observe_rate(rn, &r_prev, &rate);
            if(std::isnan(rn))
                break; // a NaN in the residual: not converged
            if(converges(rn, init, rel_tol, abs_tol))
            {
                conv = true;
                break;
            }
            if(near)
                LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2, fast));
            LAL_TRY(lal::dot_partials<float, double>(c, n, z32, res, part_g[cur ^ 1]));
            LAL_TRY(lal::update_p<float, double>(c, n, z32, p, part_g[cur ^ 1], part_g[cur]));
            cur ^= 1;
            iter++;
        }
        if(info)
        {
            info->iterations       = iter;
            info->outer_iterations = 0;
            info->converged        = conv;
            info->initial_res_norm = init;
            info->final_res_norm   = rn;
        }
        return LAL_OK;
    }
}

namespace
{
    // cg_solver::solve_precond (cg.cpp:136-223) with the Krylov recurrences in
    // FP32 and reliable updates (Sleijpen and van der Vorst, 1996) in FP64.
    // The operator product q = A p, the vectors r, p, q, z and the correction
    // y run in FP32 (A's FP32 values, the FP32 V-cycle); inner products
    // accumulate in FP64.  r is the FP64 residual scaled by 1/||r0||, so its
    // entries stay in FP32's normal range whatever b's magnitude.  When the
    // recursive residual has fallen by `delta` since the last update, or meets
    // the tolerance, the FP64 solution takes the correction (x += y ||r0||)
    // and the recursive residual is replaced by the true FP64 residual
    // b - A x; convergence is declared on that true residual only.  The
    // Krylov space is kept across updates, so the iteration count stays that
    // of the FP64 solve while every pass but the updates moves FP32 data;
    // only an update that fails to cut the true residual tenfold restarts it.
    int pcg_reliable(lal_hierarchy_t* h, const double* b, double* x, double rel_tol,
                     double abs_tol, int max_iter, int n1, int n2, double delta,
                     lal_solve_info* info)
    {
        lal_context*   c     = h->c;
        const csr_dev& A     = h->L[0].A;
        const int      n     = A.m;
        const float*   Av32  = vals<float>(A);
        double*        s     = h->s;
        double*        r64   = h->res;
        float*         r32   = h->res32;
        float*         z32   = h->z32;
        float*         p32   = h->p32;
        float*         q32   = h->q32;
        float*         y32   = h->d32;
        double*        part_pq   = h->part;
        double*        part_g[2] = {h->part + lal::PCG_BLOCKS, h->part + 2 * lal::PCG_BLOCKS};
        double*        part_r    = h->part + 3 * lal::PCG_BLOCKS;
        unsigned       seq       = 0;
        double         init      = 0.0;
        LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, A.v, x, b, r64));
        LAL_TRY(lal::norm_partials<double>(c, n, r64, part_r));
        LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM0, &seq));
        LAL_TRY(wait_scalar(h, seq, lal::S_RNORM0, &init));
        double rn   = init; // true or recursive inf-norm, unscaled
        int    iter = 0, updates = 0;
        bool   conv = converges(rn, init, rel_tol, abs_tol);
        if(!conv)
        {
            const double sigma = 1.0 / init;
            LAL_TRY(lal::scale_cast_f64_f32(c, n, r64, sigma, r32));
            LAL_HIP(hipMemsetAsync(y32, 0, sizeof(float) * n, c->stream));
            LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2));
            LAL_HIP(hipMemcpyAsync(p32, z32, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                   c->stream));
            int cur = 0;
            LAL_TRY(lal::dot_partials<float, float>(c, n, z32, r32, part_g[cur]));
            const bool   fast  = !h->opts.exact && n1 == 1 && h->levels > 0;
            work<float>& W0    = h->w<float>(0);
            double       r_upd  = init; // true residual norm at the last update
            double       r_prev = init, rate = 0.5; // last norm, observed reduction per step
            while(!(iter > max_iter))
            {
                LAL_TRY(lal::csrmv(c, A.d, n, A.n, A.nnz, 1.0, A.rp, A.ci, Av32, p32, 0.0,
                                   q32));
                LAL_TRY(lal::dot_partials<float, float>(c, n, q32, p32, part_pq));
                LAL_TRY(lal::update_xr<float, float>(c, n, p32, q32, y32, r32, part_pq,
                                                     part_g[cur], part_r, s + lal::S_ALPHA,
                                                     W0.dinv, fast ? W0.x : nullptr, nullptr));
                LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM, &seq));
                // The next V-cycle is queued before the host reads the norm,
                // unless the observed rate of reduction predicts that this step
                // crosses the update threshold: then the host decides first, so
                // that no V-cycle is spent on a residual about to be replaced.
                const double threshold = std::max(std::max(delta * r_upd, rel_tol * init),
                                                  abs_tol);
                const bool   near      = lal::near_threshold(r_prev, rate, threshold);
                if(!near)
                    LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2, fast));
                double rs = 0.0;
                LAL_TRY(wait_scalar(h, seq, lal::S_RNORM, &rs));
                rn = rs * init;
                lal::observe_rate(rn, &r_prev, &rate);
                if(std::isnan(rn))
                    break;
                const bool update = rn <= delta * r_upd || converges(rn, init, rel_tol, abs_tol);
                bool       restart = false;
                if(update)
                {
                    // x += ||r0|| y in FP64; r <- the true residual (scaled)
                    LAL_TRY(lal::axpy_f32_f64(c, n, init, y32, x));
                    LAL_HIP(hipMemsetAsync(y32, 0, sizeof(float) * n, c->stream));
                    LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, A.v, x, b,
                                              r64));
                    LAL_TRY(lal::norm_partials<double>(c, n, r64, part_r));
                    LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM, &seq));
                    LAL_TRY(lal::scale_cast_f64_f32(c, n, r64, sigma, r32));
                    LAL_TRY(wait_scalar(h, seq, lal::S_RNORM, &rn));
                    updates++;
                    // An update that did not cut the true residual tenfold
                    // means the FP32 recurrence has drifted from the FP64 system
                    // (badly conditioned operators): the kept search direction
                    // no longer belongs to the new residual, so the Krylov space
                    // restarts from it (p = z), as iterative refinement would.
                    // Healthy runs (a reduction near delta per update) never
                    // restart and keep their iterates.
                    restart = rn > 0.1 * r_upd;
                    r_upd   = rn;
                    r_prev  = rn;
                    if(std::isnan(rn))
                        break;
                    if(converges(rn, init, rel_tol, abs_tol))
                    {
                        conv = true;
                        break;
                    }
                    LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2));
                }
                else if(near)
                    LAL_TRY(precond_graphed<float>(h, r32, z32, n1, n2, fast));
                LAL_TRY(lal::dot_partials<float, float>(c, n, z32, r32, part_g[cur ^ 1]));
                if(restart)
                    LAL_HIP(hipMemcpyAsync(p32, z32, sizeof(float) * n, hipMemcpyDeviceToDevice,
                                           c->stream));
                else
                    LAL_TRY(lal::update_p<float, float>(c, n, z32, p32, part_g[cur ^ 1],
                                                        part_g[cur]));
                cur ^= 1;
                iter++;
            }
        }
        if(info)
        {
            info->iterations       = iter;
            info->outer_iterations = updates;
            info->converged        = conv;
            info->initial_res_norm = init;
            info->final_res_norm   = rn;
        }
        return LAL_OK;
    }
}

namespace lal
{
    // True when the norm r_prev, reduced once more at the observed rate, is
    // predicted to meet the threshold (or to come within 25 % of it).
    bool near_threshold(double r_prev, double rate, double threshold)
    {
        return r_prev * rate <= 1.25 * threshold;
    }

    bool near_tolerance(double r_prev, double rate, double init, double rel_tol, double abs_tol)
    {
        return near_threshold(r_prev, rate, std::max(rel_tol * init, abs_tol));
    }

    // Records the reduction of the norm over the last step, clamped to
    // [0.05, 1] (a NaN norm gives 0.05).
    void observe_rate(double rn, double* r_prev, double* rate)
    {
        *rate   = std::min(1.0, std::max(0.05, rn / *r_prev));
        *r_prev = rn;
    }

    bool converges(double res, double init, double rel_tol, double abs_tol)
    {
        if(std::isnan(res) || std::isnan(init))
            return false;
        if(res <= abs_tol)
            return true;
        if(res / init <= rel_tol)
            return true;
        return false;
    }

    // Reduces PCG_BLOCKS partials (op 0 sum, 1 max) into s[slot] and hands the
    // value to the host through the mapped note; returns the sequence number
    // to wait for.
    int post_scalar(lal_hierarchy_t* h, int op, const double* part, int slot, unsigned* seq)
    {
        *seq = ++h->seq;
        return lal::finish_partials(h->c, op, part, h->s + slot, h->d_note, h->d_seq, *seq);
    }

    // Waits until the note carries `seq`.  The host polls the mapped word
    // and checks the stream now and then so that a failed kernel turns into
    // an error instead of a hang.
    int wait_scalar(lal_hierarchy_t* h, unsigned seq, int slot, double* v)
    {
        for(long spins = 1;; spins++)
        {
            if(__atomic_load_n(h->h_seq, __ATOMIC_ACQUIRE) == seq)
            {
                *v = *reinterpret_cast<volatile double*>(h->h_note);
                return LAL_OK;
            }
            if((spins & 1023) == 0)
            {
                hipError_t e = hipStreamQuery(h->c->stream);
                if(e == hipSuccess)
                {
                    if(__atomic_load_n(h->h_seq, __ATOMIC_ACQUIRE) == seq)
                        continue;
                    // the stream has drained: read the device copy instead
                    LAL_HIP(hipMemcpy(v, h->s + slot, sizeof(double), hipMemcpyDeviceToHost));
                    return LAL_OK;
                }
                if(e != hipErrorNotReady)
                    LAL_HIP(e);
            }
        }
    }

}

extern "C" void lal_saamg_default_options(lal_saamg_options* o)
{
    if(!o)
        return;
    o->max_levels    = 20;
    o->eps           = 0.08;
    o->relax         = 2.0 / 3.0;
    o->coarse_solver = 1;
    o->exact         = 0;
    o->build_f32     = 0;
}

extern "C" void lal_solve_default_options(lal_solve_options* o)
{
    if(!o)
        return;
    o->rel_tol        = 1e-6;
    o->abs_tol        = 1e-6;
    o->max_iter       = 1000;
    o->n1             = 1;
    o->n2             = 1;
    o->inner_rel_tol  = 1e-4;
    o->inner_max_iter = 200;
    o->max_outer      = 30;
    o->smoother       = 0;
}

namespace
{
    int setup_entry(lal_handle c, int m, int n, int nnz, const int* rp, const int* ci,
                    const double* v, const lal_saamg_options* opts, lal_hierarchy* out,
                    bool unsmoothed)
    {
        LAL_REQUIRE(c != nullptr && out != nullptr && m >= 1 && n == m && nnz >= 0);
        LAL_REQUIRE(rp != nullptr && (nnz == 0 || (ci && v)));
        *out = nullptr;
        lal_saamg_options o;
        lal_saamg_default_options(&o);
        if(unsmoothed)
            o.eps = 0.001;
        if(opts)
            o = *opts;
        LAL_REQUIRE(o.max_levels >= 0 && o.eps >= 0.0);
        // validate the input pattern before any setup kernel gathers through
        // it; the analysis describes the finest operator (a copy of the same
        // pattern), so the hierarchy keeps it for level 0
        lal_csrmv_info* d0 = nullptr;
        LAL_TRY(lal::csrmv_analysis_dev(c, m, n, nnz, rp, ci, 0, &d0));
        auto* h      = new lal_hierarchy_t;
        h->L[0].A.d  = d0;
        int   s = setup(c, m, n, nnz, rp, ci, v, o, h, unsmoothed);
        if(s != LAL_OK)
        {
            (void)hipStreamSynchronize(c->stream);
            delete h;
            return s;
        }
        c->hierarchies.push_back(h);
        *out = h;
        return LAL_OK;
    }
}

namespace lal
{
    void detach_hierarchies(lal_context* c)
    {
        for(lal_hierarchy_t* h : c->hierarchies)
            h->c = nullptr;
        c->hierarchies.clear();
    }
}

extern "C" void lal_uaamg_default_options(lal_saamg_options* o)
{
    lal_saamg_default_options(o);
    if(o)
        o->eps = 0.001;
}

extern "C" int lal_uaamg_setup(lal_handle c, int m, int n, int nnz, const int* rp,
                               const int* ci, const double* v, const lal_saamg_options* opts,
                               lal_hierarchy* out)
{
    return setup_entry(c, m, n, nnz, rp, ci, v, opts, out, true);
}

extern "C" int lal_saamg_setup(lal_handle c, int m, int n, int nnz, const int* rp,
                               const int* ci, const double* v, const lal_saamg_options* opts,
                               lal_hierarchy* out)
{
    return setup_entry(c, m, n, nnz, rp, ci, v, opts, out, false);
}

extern "C" int lal_hierarchy_destroy(lal_hierarchy h)
{
    if(h)
    {
        if(h->c)
        {
            (void)hipStreamSynchronize(h->c->stream);
            auto& v = h->c->hierarchies;
            v.erase(std::remove(v.begin(), v.end(), h), v.end());
        }
        else
            (void)hipDeviceSynchronize(); // its context is gone: wait for all work
        delete h;
    }
    return LAL_OK;
}

extern "C" int lal_hierarchy_levels(lal_hierarchy h, int* total)
{
    LAL_REQUIRE(h != nullptr && total != nullptr);
    *total = h->levels;
    return LAL_OK;
}

extern "C" int lal_hierarchy_level(lal_hierarchy h, int level, int kind, int* m, int* n,
                                   int* nnz, const int** rp, const int** ci, const double** v)
{
    LAL_REQUIRE(h != nullptr && m && n && nnz && rp && ci && v);
    LAL_REQUIRE(level >= 0 && kind >= 0 && kind <= 2);
    LAL_REQUIRE(kind == 0 ? level <= h->levels : level < h->levels);
    const csr_dev& A = kind == 0 ? h->L[level].A : (kind == 1 ? h->L[level].P : h->L[level].R);
    *m   = A.m;
    *n   = A.n;
    *nnz = A.nnz;
    *rp  = A.rp;
    *ci  = A.ci;
    *v   = A.v;
    return LAL_OK;
}

extern "C" int lal_hierarchy_column_bytes(lal_hierarchy h, int level, int kind, int* bytes)
{
    LAL_REQUIRE(h != nullptr && bytes != nullptr);
    LAL_REQUIRE(level >= 0 && kind >= 0 && kind <= 2);
    LAL_REQUIRE(kind == 0 ? level <= h->levels : level < h->levels);
    const csr_dev& A = kind == 0 ? h->L[level].A : (kind == 1 ? h->L[level].P : h->L[level].R);
    *bytes           = A.d && A.d->d_c16 ? 2 : 4;
    return LAL_OK;
}

extern "C" int lal_amg_precond_f64(lal_handle c, lal_hierarchy h, const double* r, double* z,
                                   int n1, int n2)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && r != nullptr && z != nullptr);
    LAL_REQUIRE(n1 >= 0 && n2 >= 0 && c == h->c);
    return precond<double>(h, r, z, n1, n2);
}

extern "C" int lal_pcg_solve_f64(lal_handle c, lal_hierarchy h, const double* b, double* x,
                                 const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.max_iter >= 0 && o.n1 >= 0 && o.n2 >= 0);
    LAL_REQUIRE(o.smoother >= 0 && o.smoother <= 6);
    return pcg<double>(h, b, x, h->res, h->z, h->p, h->q, o.rel_tol, o.abs_tol, o.max_iter,
                       o.n1, o.n2, info, nullptr, o.smoother);
}

extern "C" int lal_pcg_solve_op_f64(lal_handle c, lal_csrmv_descr descr, int m, int nnz,
                                    const int* rp, const int* ci, const double* v,
                                    lal_hierarchy h, const double* b, double* x,
                                    const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    LAL_REQUIRE(m == h->L[0].A.m && nnz >= 0 && rp != nullptr);
    LAL_REQUIRE(nnz == 0 || (ci != nullptr && v != nullptr));
    LAL_REQUIRE(descr == nullptr || (descr->m == m && descr->nnz == nnz));
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.max_iter >= 0 && o.n1 >= 0 && o.n2 >= 0);
    LAL_REQUIRE(o.smoother >= 0 && o.smoother <= 6);
    // a non-owning view of the caller's matrix (never freed here)
    csr_dev A;
    A.m   = m;
    A.n   = m;
    A.nnz = nnz;
    A.rp  = const_cast<int*>(rp);
    A.ci  = const_cast<int*>(ci);
    A.v   = const_cast<double*>(v);
    A.d   = descr;
    return pcg<double>(h, b, x, h->res, h->z, h->p, h->q, o.rel_tol, o.abs_tol, o.max_iter,
                       o.n1, o.n2, info, &A, o.smoother);
}

extern "C" int lal_pcg_solve_mixed(lal_handle c, lal_hierarchy h, const double* b, double* x,
                                   const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    if(!h->has32)
    {
        lal::set_last_error("hierarchy was built without build_f32");
        return LAL_ERR_INVALID;
    }
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.smoother == 0); // the FP32 levels run the Jacobi V-cycle only
    LAL_REQUIRE(o.max_iter >= 0 && o.n1 >= 0 && o.n2 >= 0);
    return pcg_mixed(h, b, x, o.rel_tol, o.abs_tol, o.max_iter, o.n1, o.n2, info);
}

extern "C" int lal_pcg_solve_reliable(lal_handle c, lal_hierarchy h, const double* b,
                                      double* x, const lal_solve_options* opts,
                                      lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    if(!h->has32)
    {
        lal::set_last_error("hierarchy was built without build_f32");
        return LAL_ERR_INVALID;
    }
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.smoother == 0); // the FP32 levels run the Jacobi V-cycle only
    LAL_REQUIRE(o.max_iter >= 0 && o.n1 >= 0 && o.n2 >= 0);
    // without options the update threshold is 1e-2, which also holds the
    // iteration count on badly conditioned operators
    const double delta = opts ? o.inner_rel_tol : 1e-2;
    LAL_REQUIRE(delta > 0.0 && delta < 1.0);
    return pcg_reliable(h, b, x, o.rel_tol, o.abs_tol, o.max_iter, o.n1, o.n2, delta, info);
}

extern "C" int lal_ir_solve(lal_handle c, lal_hierarchy h, const double* b, double* x,
                            const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && b != nullptr && x != nullptr && c == h->c);
    if(!h->has32)
    {
        lal::set_last_error("hierarchy was built without build_f32");
        return LAL_ERR_INVALID;
    }
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.smoother == 0); // the FP32 levels run the Jacobi V-cycle only
    const csr_dev& A = h->L[0].A;
    const int      n = A.m;
    double*        r = h->q; // FP64 outer residual lives in the FP64 q buffer
    LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, A.v, x, b, r));
    double*  part_r = h->part + 3 * lal::PCG_BLOCKS;
    unsigned seq    = 0;
    double   init   = 0.0;
    LAL_TRY(lal::norm_partials<double>(c, n, r, part_r));
    LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM0, &seq));
    LAL_TRY(wait_scalar(h, seq, lal::S_RNORM0, &init));
    double rn    = init;
    int    outer = 0, inner_total = 0;
    bool   conv  = false;
    while(true)
    {
        if(converges(rn, init, o.rel_tol, o.abs_tol))
        {
            conv = true;
            break;
        }
        if(outer >= o.max_outer)
            break;
        // inner FP32 solve of A d = r/|r|; level 0's FP32 rhs buffer is not
        // used by the V-cycle, so it holds the inner right-hand side
        float* rhs32 = h->L[0].w32.b;
        LAL_TRY(lal::scale_cast_f64_f32(c, n, r, 1.0 / rn, rhs32));
        LAL_HIP(hipMemsetAsync(h->d32, 0, sizeof(float) * n, c->stream));
        lal_solve_info ii{};
        LAL_TRY(pcg<float>(h, rhs32, h->d32, h->res32, h->z32, h->p32, h->q32,
                           o.inner_rel_tol, 1e-30, o.inner_max_iter, o.n1, o.n2, &ii));
        inner_total += ii.iterations + 1;
        // x += |r| d in FP64, then the true FP64 residual
        LAL_TRY(lal::axpy_f32_f64(c, n, rn, h->d32, x));
        LAL_TRY(lal::residual_f64(c, A.d, n, A.n, A.nnz, A.rp, A.ci, A.v, x, b, r));
        LAL_TRY(lal::norm_partials<double>(c, n, r, part_r));
        LAL_TRY(post_scalar(h, 1, part_r, lal::S_RNORM, &seq));
        LAL_TRY(wait_scalar(h, seq, lal::S_RNORM, &rn));
        outer++;
        if(std::isnan(rn))
            break;
    }
    if(info)
    {
        info->iterations       = inner_total;
        info->outer_iterations = outer;
        info->converged        = conv;
        info->initial_res_norm = init;
        info->final_res_norm   = rn;
    }
    return LAL_OK;
}

extern "C" int lal_adjoint_solve(lal_handle c, lal_hierarchy h, const double* g, double* lambda,
                                 int mixed, const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_REQUIRE(c != nullptr && h != nullptr && g != nullptr && lambda != nullptr && c == h->c);
    LAL_REQUIRE(mixed >= 0 && mixed <= 3);
    LAL_HIP(hipMemsetAsync(lambda, 0, sizeof(double) * h->L[0].A.m, c->stream));
    if(mixed == 1)
        return lal_ir_solve(c, h, g, lambda, opts, info);
    if(mixed == 2)
        return lal_pcg_solve_mixed(c, h, g, lambda, opts, info);
    if(mixed == 3)
        return lal_pcg_solve_reliable(c, h, g, lambda, opts, info);
    return lal_pcg_solve_f64(c, h, g, lambda, opts, info);
}
