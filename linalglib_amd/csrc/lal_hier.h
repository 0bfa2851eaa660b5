// SA-AMG hierarchy as the device code sees it (shared by lal_hierarchy.hip
// and lal_cycles.hip).
#pragma once

#include "lal_coarse.h"
#include "lal_vcycle.h"

#include <algorithm>
#include <vector>

struct lal_hierarchy_t;

namespace lal
{
    struct csr_dev
    {
        int             m = 0, n = 0, nnz = 0;
        int*            rp  = nullptr;
        int*            ci  = nullptr;
        double*         v   = nullptr;
        float*          v32 = nullptr;
        lal_csrmv_info* d   = nullptr;

        void free_all()
        {
            (void)hipFree(rp);
            (void)hipFree(ci);
            (void)hipFree(v);
            (void)hipFree(v32);
            lal_csrmv_descr_destroy(d);
            rp = ci = nullptr;
            v       = nullptr;
            v32     = nullptr;
            d       = nullptr;
        }
    };

    template <typename T>
    struct work
    {
        T* x    = nullptr;
        T* t    = nullptr;
        T* b    = nullptr;
        T* r    = nullptr;
        T* diag = nullptr; // a_ii (exact path)
        T* dinv = nullptr; // 1/a_ii (fast path)
    };

    // Level sets of the Gauss-Seidel sweeps of one operator: the rows of set k
    // depend only on rows of sets < k, so a set is one launch and each row is
    // updated in the order the sequential sweep would use (forward, and
    // backward for the second half of SSOR).  Operators whose sets are thin
    // are swept by one workgroup from a copy of the matrix with its rows in
    // set order (prp/pci/pv; a diagonal entry of column c is stored as
    // -1 - c), so that a set's entries are contiguous.
    struct gs_sched
    {
        int*             perm  = nullptr; // rows grouped by set
        std::vector<int> off;             // set k: perm[off[k] .. off[k+1])
        int*             off_d = nullptr; // the same offsets on the device
        int*             prp   = nullptr; // row pointer of the set-ordered copy
        int*             pci   = nullptr;
        double*          pv    = nullptr;
        double*          pdiag = nullptr; // a_jj per row of the copy (last diagonal entry)
        int*             tiles = nullptr; // 4 ints per tile of the one-workgroup sweep
        int              ntiles = 0;
    };

    // Multicolour Gauss-Seidel schedule of one operator (lal_mcgs.hip): rows
    // grouped by colour (colour k: perm[off[k] .. off[k+1])) and a copy of the
    // matrix in that row order without its diagonal, plus 1/a_ii per row.
    struct mc_sched
    {
        bool             built = false;
        int              ncol  = 0;
        int              lanes = 1;     // lanes per row
        bool             wg    = false; // every colour in one workgroup
        std::vector<int> off;
        int*             off_d = nullptr;
        int*             perm  = nullptr;
        int*             prp   = nullptr;
        int*             pci   = nullptr;
        double*          pv    = nullptr;
        double*          pdinv = nullptr;
        int*             plan    = nullptr; // one-workgroup sweep: 8 ints per chunk
        std::vector<int> plan_h;            // the same plan on the host
        int              nchunks = 0;       // chunks of one sweep direction
        int              nnz     = 0;       // entries of the copy
        int              wg_nt   = 0;       // lanes of the one-workgroup sweep
        int              wg_lg   = 0;       // log2 of its lanes per row
    };

    struct level_t
    {
        csr_dev      A, P, R;
        work<double> w64;
        work<float>  w32;
        gs_sched     gs_fwd, gs_bwd;
        mc_sched     mc;
    };

    // z = one V-cycle from zero on r with smoother 0-6 of the AMG cycles
    // (lal_cycles.hip: n1 pre and n2 post sweeps per level, the reference's
    // coarsest solve).
    int vcycle_smoothed(lal_hierarchy_t* h, const double* r, double* z, int n1, int n2,
                        int smoother);
    // Builds the multicolour schedule of A (host colouring, device copy).
    int build_mc(lal_context* c, const csr_dev& A, mc_sched& s);
    // One multicolour Gauss-Seidel sweep on x in place (forward colours, then
    // backward when sym).
    int mc_sweep(lal_context* c, const mc_sched& s, int m, const double* b, double* x, bool sym);

    inline int alloc_bytes(void** p, size_t bytes)
    {
        *p = nullptr;
        if(bytes == 0)
            return LAL_OK;
        LAL_HIP(hipMalloc(p, bytes));
        return LAL_OK;
    }

    template <typename T>
    int alloc_n(T** p, size_t n)
    {
        return alloc_bytes((void**)p, sizeof(T) * std::max<size_t>(n, 1));
    }
}

struct lal_hierarchy_t
{
    lal_context*      c = nullptr;
    lal_saamg_options opts{};
    int               levels = 0;
    lal::level_t      L[21];
    // coarsest level
    int     coarse_kind = 0; // 0 scalar, 1 Gauss-Seidel, 2 dense inverse
    double* coarse_inv  = nullptr;
    // PCG workspaces and device scalars
    double* res = nullptr;
    double* z   = nullptr;
    double* p   = nullptr;
    double* q   = nullptr;
    float*  res32 = nullptr;
    float*  z32   = nullptr;
    float*  p32   = nullptr;
    float*  q32   = nullptr;
    float*  d32   = nullptr;
    double* s     = nullptr; // LAL_SOLVER_SLOTS scalars
    double* part  = nullptr; // 4 x PCG_BLOCKS reduction partials
    // residual norms reach the host through mapped pinned memory: a value
    // and a sequence number the host polls
    double*   h_note     = nullptr;
    double*   d_note     = nullptr;
    unsigned* h_seq      = nullptr;
    unsigned* d_seq      = nullptr;
    unsigned  seq        = 0;
    bool      has32      = false;
    // coarse tail: levels tail_first..levels run in one workgroup out of LDS
    // (fast path); one packed device image of their arrays per precision
    int            tail_first = -1;
    int*           tail_img64 = nullptr;
    int*           tail_img32 = nullptr;
    lal::tail_desc tail64{};
    lal::tail_desc tail32{};
    // captured cycles (lal::captured): the multicolour smoothers' cycles
    // (kind = smoother 5 or 6) and the PCG preconditioner's V-cycle (kind
    // lal::KIND_PRECOND + 2 xinit + FP32), one per (x, b, sweeps, cycle, kind)
    struct cycle_graph
    {
        const void*    x;
        const void*    b;
        int            n1, n2, n3, cycle, smoother;
        bool           failed;
        hipGraphExec_t exec;
    };
    std::vector<cycle_graph> cycle_graphs;
    hipStream_t              capture_stream = nullptr; // where they are captured

    ~lal_hierarchy_t()
    {
        for(auto& g : cycle_graphs)
            if(g.exec)
                (void)hipGraphExecDestroy(g.exec);
        if(capture_stream)
            (void)hipStreamDestroy(capture_stream);
        for(int l = 0; l <= 20; l++)
        {
            L[l].A.free_all();
            L[l].P.free_all();
            L[l].R.free_all();
            for(double* ptr : {L[l].w64.x, L[l].w64.t, L[l].w64.b, L[l].w64.r, L[l].w64.diag,
                               L[l].w64.dinv})
                (void)hipFree(ptr);
            for(float* ptr : {L[l].w32.x, L[l].w32.t, L[l].w32.b, L[l].w32.r, L[l].w32.diag,
                              L[l].w32.dinv})
                (void)hipFree(ptr);
            for(lal::gs_sched* g : {&L[l].gs_fwd, &L[l].gs_bwd})
                for(void* ptr : {(void*)g->perm, (void*)g->off_d, (void*)g->prp, (void*)g->pci,
                                 (void*)g->pv, (void*)g->pdiag, (void*)g->tiles})
                    (void)hipFree(ptr);
            for(void* ptr : {(void*)L[l].mc.off_d, (void*)L[l].mc.perm, (void*)L[l].mc.prp,
                             (void*)L[l].mc.pci, (void*)L[l].mc.pv, (void*)L[l].mc.pdinv,
                             (void*)L[l].mc.plan})
                (void)hipFree(ptr);
        }
        for(void* ptr : {(void*)coarse_inv, (void*)res, (void*)z, (void*)p, (void*)q,
                         (void*)res32, (void*)z32, (void*)p32, (void*)q32, (void*)d32, (void*)s,
                         (void*)part, (void*)tail_img64, (void*)tail_img32})
            (void)hipFree(ptr);
        if(h_note)
            (void)hipHostFree(h_note);
    }

    template <typename T>
    lal::work<T>& w(int l);
};

template <>
inline lal::work<double>& lal_hierarchy_t::w<double>(int l)
{
    return L[l].w64;
}
template <>
inline lal::work<float>& lal_hierarchy_t::w<float>(int l)
{
    return L[l].w32;
}


namespace lal
{
    // iter_control::residual_converges (iter_control.cpp:38-78)
    bool converges(double res, double init, double rel_tol, double abs_tol);
    // Convergence prediction for the solver loops, which queue the next
    // preconditioner application before the host reads a residual norm unless
    // the step is predicted to converge (or to cross a threshold): true when
    // r_prev, reduced once more at the observed rate, meets the threshold or
    // comes within 25 % of it.
    bool near_threshold(double r_prev, double rate, double threshold);
    bool near_tolerance(double r_prev, double rate, double init, double rel_tol, double abs_tol);
    // Records the reduction over the last step, clamped to [0.05, 1] (a NaN
    // norm gives 0.05), and the new norm.
    void observe_rate(double rn, double* r_prev, double* rate);
    // Reduces PCG_BLOCKS partials (op 0 sum, 1 max) into the solver slot and
    // hands the value to the host through the hierarchy's mapped note;
    // returns the sequence number to wait for.
    int post_scalar(lal_hierarchy_t* h, int op, const double* part, int slot, unsigned* seq);
    // Waits for the note carrying `seq` (checks the stream now and then).
    int wait_scalar(lal_hierarchy_t* h, unsigned seq, int slot, double* v);

    constexpr int KIND_PRECOND = 100;

    // Runs body() -- launches on h->c->stream only: no allocation, memset or
    // host synchronisation -- eagerly on the first use of its key, captures it
    // as a HIP graph on the hierarchy's own non-blocking stream on the second
    // (the caller's stream may be the null stream, which cannot be captured;
    // the context's stream points at the capture stream meanwhile) and
    // replays the graph on the caller's stream from then on.  A capture that
    // fails leaves the key eager.
    template <typename Body>
    int captured(lal_hierarchy_t* h, const void* x, const void* b, int n1, int n2, int n3,
                 int cycle, int kind, Body&& body)
    {
        lal_context*                  c = h->c;
        lal_hierarchy_t::cycle_graph* g = nullptr;
        for(auto& e : h->cycle_graphs)
            if(e.x == x && e.b == b && e.n1 == n1 && e.n2 == n2 && e.n3 == n3 &&
               e.cycle == cycle && e.smoother == kind)
                g = &e;
        if(g == nullptr)
        {
            LAL_TRY(body());
            h->cycle_graphs.push_back({x, b, n1, n2, n3, cycle, kind, false, nullptr});
            return LAL_OK;
        }
        if(g->exec)
        {
            LAL_HIP(hipGraphLaunch(g->exec, c->stream));
            return LAL_OK;
        }
        if(g->failed)
            return body();
        if(h->capture_stream == nullptr &&
           hipStreamCreateWithFlags(&h->capture_stream, hipStreamNonBlocking) != hipSuccess)
        {
            (void)hipGetLastError();
            h->capture_stream = nullptr;
            g->failed         = true;
            return body();
        }
        hipGraph_t        graph = nullptr;
        hipGraphExec_t    exec  = nullptr;
        const hipStream_t user  = c->stream;
        if(hipStreamBeginCapture(h->capture_stream, hipStreamCaptureModeThreadLocal) ==
           hipSuccess)
        {
            c->stream          = h->capture_stream;
            const int        st = body();
            c->stream          = user;
            const hipError_t e  = hipStreamEndCapture(h->capture_stream, &graph);
            if(st == LAL_OK && e == hipSuccess && graph != nullptr &&
               hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess)
                g->exec = exec;
            if(graph)
                (void)hipGraphDestroy(graph);
        }
        (void)hipGetLastError();
        if(g->exec == nullptr)
        {
            g->failed = true;
            return body();
        }
        LAL_HIP(hipGraphLaunch(g->exec, c->stream));
        return LAL_OK;
    }
}
