// BiCGSTAB and restarted GMRES on the device with no, Jacobi or SA-AMG
// preconditioning, and the Jacobi preconditioner itself.
//
//   jacobi_precond      iterative_solvers/preconditioner/preconditioner.cpp:41-58,
//                       backend/host/host_math.cpp:53-64 (z = r / a_ii)
//   bicgstab_solver     iterative_solvers/krylov/bicgstab.cpp:64-231, including
//                       the breakdown branch's x += alpha p
//   gmres_solver        iterative_solvers/krylov/gmres.cpp:46-530: restarted,
//                       left-preconditioned, modified Gram-Schmidt Arnoldi,
//                       Givens rotations, convergence on the preconditioned
//                       residual 2-norm, restart cycles counted as iterations.
//                       gmres.cpp drops the preconditioned Arnoldi vector (it
//                       lands in a temporary) and never sizes its residual
//                       vectors; here q_k = M^-1 A q_(k-1) is kept and every
//                       work vector is sized.
//
// The vectors stay on the device.  BiCGSTAB keeps its scalars on the device
// (inner products as block partials summed by the kernel that uses them) and
// reads one pair of numbers per iteration, the residual norm and the
// breakdown flag; GMRES keeps its Gram-Schmidt inner products on the device
// and copies one Hessenberg column per Arnoldi step to the host, where the
// Givens rotations and the small triangular solve run.  Work vectors are
// stream-ordered allocations of the call.
#include "lal_hier.h"

#include <cmath>
#include <new>
#include <vector>

namespace
{
    __global__ void __launch_bounds__(256)
        div_kernel(int n, const double* __restrict__ x, const double* __restrict__ d,
                   double* __restrict__ y)
    {
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            y[i] = x[i] / d[i];
    }

    __global__ void __launch_bounds__(256)
        div_scalar_kernel(int n, const double* __restrict__ x, double s, double* __restrict__ y)
    {
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            y[i] = x[i] / s;
    }

    int div_vec(lal_context* c, int n, const double* x, const double* d, double* y)
    {
        if(n == 0)
            return LAL_OK;
        div_kernel<<<lal::grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, x, d, y);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    int div_scalar(lal_context* c, int n, const double* x, double s, double* y)
    {
        if(n == 0)
            return LAL_OK;
        div_scalar_kernel<<<lal::grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, x, s, y);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    // Modified Gram-Schmidt on the device: the inner products stay on the
    // device as lal::PCG_BLOCKS partials that every block of the next kernel
    // sums in the same order, so a whole Arnoldi step needs one copy of its
    // Hessenberg column to the host.
    constexpr int KT = 256;
    constexpr int KB = lal::PCG_BLOCKS;

    __device__ double block_sum(double v)
    {
        __shared__ double sh[KT / 64];
        v = lal::wave_sum(v);
        __syncthreads();
        if((threadIdx.x & 63) == 0)
            sh[threadIdx.x >> 6] = v;
        __syncthreads();
        double t = sh[0];
#pragma unroll
        for(int w = 1; w < KT / 64; w++)
            t = t + sh[w];
        return t;
    }

    __device__ double sum_parts(const double* __restrict__ part)
    {
        double v = 0.0;
        for(int k = threadIdx.x; k < KB; k += KT)
            v += part[k];
        return block_sum(v);
    }

    // h = <q_i, q_k> from part_in (written to *h_out by block 0);
    // q_k = q_k - h q_i; part_out = partials of <q_next, q_k>, or of
    // <q_k, q_k> when q_next is null (the norm that ends the step).
    __global__ void __launch_bounds__(KT)
        mgs_kernel(int n, const double* __restrict__ part_in, const double* __restrict__ qi,
                   const double* __restrict__ qnext, double* __restrict__ qk,
                   double* __restrict__ part_out, double* __restrict__ h_out)
    {
        const double h = sum_parts(part_in);
        if(blockIdx.x == 0 && threadIdx.x == 0)
            *h_out = h;
        double acc = 0.0;
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
        {
            const double q = qk[i] - h * qi[i];
            qk[i]          = q;
            acc            = acc + (qnext ? qnext[i] : q) * q;
        }
        const double t = block_sum(acc);
        if(threadIdx.x == 0)
            part_out[blockIdx.x] = t;
    }

    // h = sqrt(<q_k, q_k>) from part (to *h_out); q_k = q_k / h.
    __global__ void __launch_bounds__(KT)
        normalize_kernel(int n, const double* __restrict__ part, double* __restrict__ qk,
                         double* __restrict__ h_out)
    {
        const double h = sqrt(sum_parts(part));
        if(blockIdx.x == 0 && threadIdx.x == 0)
            *h_out = h;
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
            qk[i] = qk[i] / h;
    }

    // BiCGSTAB with its scalars on the device.  sc[] holds alpha, omega, the
    // breakdown flag and the residual norm; the inner products are KB partials
    // that every block of the consuming kernel sums in the same order.
    enum
    {
        SC_ALPHA = 0,
        SC_OMEGA = 1,
        SC_FLAG  = 2,
        SC_RNORM = 3,
        SC_COUNT = 4
    };

    // partials of <t, r> and <t, t> in one pass
    __global__ void __launch_bounds__(KT)
        dot2_kernel(int n, const double* __restrict__ t, const double* __restrict__ r,
                    double* __restrict__ part1, double* __restrict__ part2)
    {
        double a = 0.0, c = 0.0;
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
        {
            const double ti = t[i];
            a               = a + ti * r[i];
            c               = c + ti * ti;
        }
        const double sa = block_sum(a);
        const double sc = block_sum(c);
        if(threadIdx.x == 0)
        {
            part1[blockIdx.x] = sa;
            part2[blockIdx.x] = sc;
        }
    }

    // alpha = rho / <r0, A d>; r = -alpha (A d) + r
    __global__ void __launch_bounds__(KT)
        bicg_alpha_kernel(int n, const double* __restrict__ part_rho,
                          const double* __restrict__ part_den, const double* __restrict__ ad,
                          double* __restrict__ r, double* __restrict__ sc)
    {
        const double rho   = sum_parts(part_rho);
        const double den   = sum_parts(part_den);
        const double alpha = rho / den;
        if(blockIdx.x == 0 && threadIdx.x == 0)
            sc[SC_ALPHA] = alpha;
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
            r[i] = -1.0 * alpha * ad[i] + r[i];
    }

    // omega = <t, r> / <t, t>; unless either is exactly zero (then x += alpha p
    // and the breakdown flag is raised): x = alpha d + omega s + x and
    // r = -omega t + r
    __global__ void __launch_bounds__(KT)
        bicg_omega_kernel(int n, const double* __restrict__ part1,
                          const double* __restrict__ part2, const double* __restrict__ d,
                          const double* __restrict__ s, const double* __restrict__ t,
                          const double* __restrict__ p, double* __restrict__ x,
                          double* __restrict__ r, double* __restrict__ sc)
    {
        const double w1    = sum_parts(part1);
        const double w2    = sum_parts(part2);
        const double alpha = sc[SC_ALPHA];
        const bool   brk   = w1 == 0.0 || w2 == 0.0;
        const double omega = brk ? 0.0 : w1 / w2;
        if(blockIdx.x == 0 && threadIdx.x == 0)
        {
            sc[SC_OMEGA] = omega;
            sc[SC_FLAG]  = brk ? 1.0 : 0.0;
        }
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
        {
            if(brk)
                x[i] = alpha * p[i] + x[i];
            else
            {
                x[i] = alpha * d[i] + omega * s[i] + 1.0 * x[i];
                r[i] = -1.0 * omega * t[i] + r[i];
            }
        }
    }

    // beta = (rho_new / rho_old) (alpha / omega); p = 1 r - beta omega (A d) + beta p
    __global__ void __launch_bounds__(KT)
        bicg_p_kernel(int n, const double* __restrict__ part_new,
                      const double* __restrict__ part_old, const double* __restrict__ r,
                      const double* __restrict__ ad, double* __restrict__ p,
                      const double* __restrict__ sc)
    {
        const double rho_new = sum_parts(part_new);
        const double rho_old = sum_parts(part_old);
        const double alpha   = sc[SC_ALPHA];
        const double omega   = sc[SC_OMEGA];
        const double beta    = (rho_new / rho_old) * (alpha / omega);
        const double g       = -1.0 * beta * omega;
        for(int i = blockIdx.x * KT + threadIdx.x; i < n; i += KB * KT)
            p[i] = 1.0 * r[i] + g * ad[i] + beta * p[i];
    }

    // Stream-ordered work vectors released when the solver returns, and an
    // event that marks a scalar read-back on the stream.
    struct workspace
    {
        lal_context*         c;
        std::vector<double*> bufs;
        hipEvent_t           ev = nullptr;
        explicit workspace(lal_context* c_) : c(c_) {}
        int make_event()
        {
            LAL_HIP(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
            return LAL_OK;
        }
        int get(double** p, size_t n)
        {
            *p = nullptr;
            LAL_HIP(hipMallocAsync((void**)p, sizeof(double) * std::max<size_t>(n, 1), c->stream));
            bufs.push_back(*p);
            return LAL_OK;
        }
        ~workspace()
        {
            for(double* p : bufs)
                (void)hipFreeAsync(p, c->stream);
            if(ev)
                (void)hipEventDestroy(ev);
        }
    };

    struct precond_t
    {
        int           kind = 0; // 0 none, 1 Jacobi, 2 SA-AMG V-cycle
        lal_hierarchy h    = nullptr;
        const double* diag = nullptr;
        int           n1 = 1, n2 = 1;

        int apply(lal_context* c, int n, const double* r, double* z) const
        {
            if(kind == 1)
                return div_vec(c, n, r, diag, z);
            if(kind == 2)
                return lal_amg_precond_f64(c, h, r, z, n1, n2);
            return lal_memcpy_d2d(c, z, r, sizeof(double) * n);
        }
    };

    struct op_t
    {
        lal_csrmv_descr d;
        int             m, nnz;
        const int*      rp;
        const int*      ci;
        const double*   v;
        int             spmv(lal_context* c, const double* x, double* y) const
        {
            return lal_csrmv_f64(c, d, m, m, nnz, 1.0, rp, ci, v, x, 0.0, y);
        }
        int residual(lal_context* c, const double* x, const double* b, double* r) const
        {
            return lal_residual_f64(c, d, m, m, nnz, rp, ci, v, x, b, r);
        }
    };

    int setup_precond(lal_context* c, workspace& ws, const op_t& A, int kind, lal_hierarchy h,
                      const lal_solve_options& o, precond_t* P)
    {
        P->kind = kind;
        P->h    = h;
        P->n1   = o.n1;
        P->n2   = o.n2;
        if(kind == 1)
        {
            double* d = nullptr;
            LAL_TRY(ws.get(&d, A.m));
            LAL_TRY(lal_fill_f64(c, d, A.m, 0.0));
            LAL_TRY(lal_diagonal_f64(c, A.m, A.rp, A.ci, A.v, d));
            P->diag = d;
        }
        return LAL_OK;
    }

    int check_args(lal_context* c, int m, int nnz, const int* rp, const double* b, double* x,
                   int precond, lal_hierarchy h)
    {
        LAL_REQUIRE(c != nullptr && m >= 1 && nnz >= 0 && rp != nullptr && b && x);
        LAL_REQUIRE(precond >= 0 && precond <= 2);
        LAL_REQUIRE(precond != 2 || h != nullptr);
        if(precond == 2)
        {
            // the V-cycle reads and writes vectors of the finest level's size
            int           hm = 0, hn = 0, hnnz = 0;
            const int *   hrp = nullptr, *hci = nullptr;
            const double* hv = nullptr;
            LAL_TRY(lal_hierarchy_level(h, 0, 0, &hm, &hn, &hnnz, &hrp, &hci, &hv));
            LAL_REQUIRE(hm == m && hnnz == nnz);
        }
        return LAL_OK;
    }
}

extern "C" int lal_jacobi_solve_f64(lal_handle c, int n, const double* rhs, const double* diag,
                                    double* x)
{
    LAL_REQUIRE(c != nullptr && n >= 0 && (n == 0 || (rhs && diag && x)));
    return div_vec(c, n, rhs, diag, x);
}

extern "C" int lal_bicgstab_solve(lal_handle c, lal_csrmv_descr d, int m, int nnz,
                                  const int* rp, const int* ci, const double* val,
                                  const double* b, double* x, int precond, lal_hierarchy h,
                                  const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_TRY(check_args(c, m, nnz, rp, b, x, precond, h));
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.max_iter >= 0);
    const op_t A{d, m, nnz, rp, ci, val};
    workspace  ws(c);
    double *   r, *r0, *p, *vv, *t, *z, *q;
    for(double** buf : {&r, &r0, &p, &vv, &t, &z, &q})
        LAL_TRY(ws.get(buf, m));
    precond_t P;
    LAL_TRY(setup_precond(c, ws, A, precond, h, o, &P));
    const bool pre = precond != 0;
    double *   sc, *part;
    LAL_TRY(ws.get(&sc, SC_COUNT));
    LAL_TRY(ws.get(&part, 5 * (size_t)KB));
    double* part_rho[2] = {part, part + KB};
    double* part_den    = part + 2 * KB;
    double* part_w1     = part + 3 * KB;
    double* part_w2     = part + 4 * KB;

    LAL_TRY(A.residual(c, x, b, r));
    double init = 0.0, rn = 0.0;
    LAL_TRY(lal_norm_inf_f64(c, m, r, &init));
    rn = init;
    LAL_TRY(lal_memcpy_d2d(c, r0, r, sizeof(double) * m));
    int cur = 0;
    LAL_TRY(lal::dot_partials<double, double>(c, m, r0, r, part_rho[cur]));
    LAL_TRY(lal_memcpy_d2d(c, p, r, sizeof(double) * m));
    if(pre)
        LAL_TRY(P.apply(c, m, r, z));
    int  iter = 0;
    bool conv = false;
    // pinned read-back of {flag, norm}: slots 8 and 9 of the context's mirror
    // (slot 0 belongs to the blocking reductions), so the copy is asynchronous
    // and the next direction's kernels queue behind it
    double* hs = c->h_slots + 8;
    LAL_TRY(ws.make_event());
    double r_prev = init, rate = 0.5; // last norm, observed reduction per step
    // rho_new = <r0, r>; p = r - beta omega (A d) + beta p; z = M^-1 p
    auto next_direction = [&](const double* Ad) -> int {
        LAL_TRY(lal::dot_partials<double, double>(c, m, r0, r, part_rho[cur ^ 1]));
        bicg_p_kernel<<<KB, KT, 0, c->stream>>>(m, part_rho[cur ^ 1], part_rho[cur], r, Ad, p,
                                                sc);
        LAL_LAUNCHED();
        if(pre)
            LAL_TRY(P.apply(c, m, p, z));
        return LAL_OK;
    };
    while(!(iter > o.max_iter))
    {
        const double* dir = pre ? z : p;
        double*       Ad  = pre ? q : vv;
        LAL_TRY(A.spmv(c, dir, Ad));
        LAL_TRY(lal::dot_partials<double, double>(c, m, r0, Ad, part_den));
        bicg_alpha_kernel<<<KB, KT, 0, c->stream>>>(m, part_rho[cur], part_den, Ad, r, sc);
        LAL_LAUNCHED();
        if(pre)
            LAL_TRY(P.apply(c, m, r, vv));
        const double* sv = pre ? vv : r;
        LAL_TRY(A.spmv(c, sv, t));
        dot2_kernel<<<KB, KT, 0, c->stream>>>(m, t, r, part_w1, part_w2);
        LAL_LAUNCHED();
        bicg_omega_kernel<<<KB, KT, 0, c->stream>>>(m, part_w1, part_w2, dir, sv, t, p, x, r,
                                                    sc);
        LAL_LAUNCHED();
        LAL_TRY(lal::norm_inf_f64_dev(c, m, r, sc + SC_RNORM, 0));
        LAL_HIP(hipMemcpyAsync(hs, sc + SC_FLAG, 2 * sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipEventRecord(ws.ev, c->stream));
        // The next direction depends on r only: it is queued before the host
        // waits for the read-back, so the GPU keeps working while the host
        // tests the norm -- unless the observed rate of reduction predicts
        // that this step converges, when the host decides first.
        const bool near = lal::near_tolerance(r_prev, rate, init, o.rel_tol, o.abs_tol);
        if(!near)
            LAL_TRY(next_direction(Ad));
        LAL_HIP(hipEventSynchronize(ws.ev));
        if(hs[0] != 0.0)
            break; // breakdown: x += alpha p was applied, as bicgstab.cpp does
        rn = hs[1];
        lal::observe_rate(rn, &r_prev, &rate);
        if(lal::converges(rn, init, o.rel_tol, o.abs_tol))
        {
            conv = true;
            break;
        }
        if(near)
            LAL_TRY(next_direction(Ad));
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

extern "C" int lal_gmres_solve(lal_handle c, lal_csrmv_descr d, int m, int nnz, const int* rp,
                               const int* ci, const double* val, const double* b, double* x,
                               int restart, int precond, lal_hierarchy h,
                               const lal_solve_options* opts, lal_solve_info* info)
{
    LAL_TRY(check_args(c, m, nnz, rp, b, x, precond, h));
    LAL_REQUIRE(restart >= 1);
    lal_solve_options o;
    lal_solve_default_options(&o);
    if(opts)
        o = *opts;
    LAL_REQUIRE(o.max_iter >= 0);
    restart = std::max(1, std::min(restart, m - 1));
    const int  ld = restart + 1;
    const op_t A{d, m, nnz, rp, ci, val};
    workspace  ws(c);
    double *   Q, *w, *z, *res, *dH, *part;
    LAL_TRY(ws.get(&Q, (size_t)m * ld));
    for(double** buf : {&w, &z, &res})
        LAL_TRY(ws.get(buf, m));
    LAL_TRY(ws.get(&dH, (size_t)ld * restart));
    LAL_TRY(ws.get(&part, 2 * (size_t)KB));
    precond_t P;
    LAL_TRY(setup_precond(c, ws, A, precond, h, o, &P));
    // the host Hessenberg matrix is (restart + 1) x restart: refuse what the
    // host cannot hold instead of letting an allocation failure escape the C ABI
    std::vector<double> H, cs, sn, y;
    try
    {
        H.assign((size_t)ld * restart, 0.0);
        cs.assign(restart, 0.0);
        sn.assign(restart, 0.0);
        y.assign(ld, 0.0);
    }
    catch(const std::bad_alloc&)
    {
        lal::set_last_error("gmres: restart too large for the host Hessenberg matrix");
        return LAL_ERR_OOM;
    }
    auto col = [&](int k) { return Q + (size_t)k * m; };
    // q_(k+1) = M^-1 A q_k (before its orthogonalisation)
    auto arnoldi_product = [&](int k) -> int {
        LAL_TRY(A.spmv(c, col(k), w));
        return P.apply(c, m, w, col(k + 1));
    };
    LAL_TRY(ws.make_event());

    LAL_TRY(A.residual(c, x, b, res));
    LAL_TRY(P.apply(c, m, res, z));
    double vv = 0.0;
    LAL_TRY(lal_dot_f64(c, m, z, z, &vv));
    double rn    = std::sqrt(vv);
    double init  = rn;
    int    iter  = 0;
    int    steps = 0;
    bool   conv = lal::converges(rn, init, o.rel_tol, o.abs_tol);
    if(!conv)
    {
        LAL_TRY(div_scalar(c, m, z, rn, col(0)));
        std::fill(y.begin(), y.end(), 0.0);
        y[0] = rn;
        while(!(iter > o.max_iter))
        {
            LAL_TRY(arnoldi_product(0));
            double r_prev = y[0], rate = 0.5; // residual estimate, reduction per step
            int    k      = 1;
            for(k = 1; k < restart + 1; k++)
            {
                double* qk = col(k);
                steps++;
                // h_ik = <q_i, q_k>, q_k -= h_ik q_i for i < k, in order, then
                // h_k,k-1 = ||q_k|| and q_k /= h_k,k-1, all on the device
                double* hcol = dH + (size_t)(k - 1) * ld;
                LAL_TRY(lal::dot_partials<double, double>(c, m, col(0), qk, part));
                for(int i = 0; i < k; i++)
                {
                    const double* pin  = part + (size_t)(i & 1) * KB;
                    double*       pout = part + (size_t)((i + 1) & 1) * KB;
                    mgs_kernel<<<KB, KT, 0, c->stream>>>(m, pin, col(i),
                                                         i + 1 < k ? col(i + 1) : nullptr, qk,
                                                         pout, hcol + i);
                    LAL_LAUNCHED();
                }
                normalize_kernel<<<KB, KT, 0, c->stream>>>(m, part + (size_t)(k & 1) * KB, qk,
                                                           hcol + k);
                LAL_LAUNCHED();
                LAL_HIP(hipMemcpyAsync(&H[(size_t)(k - 1) * ld], hcol, sizeof(double) * (k + 1),
                                       hipMemcpyDeviceToHost, c->stream));
                LAL_HIP(hipEventRecord(ws.ev, c->stream));
                // the next Arnoldi product needs only the normalised q_k: it
                // is queued before the host waits for the column, unless this
                // is the cycle's last step or the residual estimate is
                // predicted to converge here
                const bool ahead =
                    k < restart && !lal::near_tolerance(r_prev, rate, init, o.rel_tol, o.abs_tol);
                if(ahead)
                    LAL_TRY(arnoldi_product(k));
                LAL_HIP(hipEventSynchronize(ws.ev));
                for(int i = 0; i < k - 1; i++)
                {
                    const double t1 = H[i + (k - 1) * ld], t2 = H[i + 1 + (k - 1) * ld];
                    H[i + (k - 1) * ld]     = cs[i] * t1 + sn[i] * t2;
                    H[i + 1 + (k - 1) * ld] = -sn[i] * t1 + cs[i] * t2;
                }
                const double xi = H[k - 1 + (k - 1) * ld], xj = H[k + (k - 1) * ld];
                if(xi == 0.0)
                {
                    cs[k - 1] = 0.0;
                    sn[k - 1] = 1.0;
                }
                else if(xj == 0.0)
                {
                    cs[k - 1] = 1.0;
                    sn[k - 1] = 0.0;
                }
                else if(std::fabs(xi) > std::fabs(xj))
                {
                    const double t = xj / xi;
                    cs[k - 1]      = 1.0 / std::sqrt(1.0 + t * t);
                    sn[k - 1]      = t * cs[k - 1];
                }
                else
                {
                    const double t = xi / xj;
                    sn[k - 1]      = 1.0 / std::sqrt(1.0 + t * t);
                    cs[k - 1]      = t * sn[k - 1];
                }
                {
                    const double t1 = H[k - 1 + (k - 1) * ld], t2 = H[k + (k - 1) * ld];
                    H[k - 1 + (k - 1) * ld] = cs[k - 1] * t1 + sn[k - 1] * t2;
                    H[k + (k - 1) * ld]     = -sn[k - 1] * t1 + cs[k - 1] * t2;
                }
                y[k]     = -sn[k - 1] * y[k - 1];
                y[k - 1] = cs[k - 1] * y[k - 1];
                lal::observe_rate(std::fabs(y[k]), &r_prev, &rate);
                if(lal::converges(std::fabs(y[k]), init, o.rel_tol, o.abs_tol) || k == restart)
                    break;
                if(!ahead)
                    LAL_TRY(arnoldi_product(k));
            }
            for(int i = k - 1; i >= 0; i--)
            {
                for(int j = i + 1; j < k; j++)
                    y[i] = y[i] - H[i + ld * j] * y[j];
                y[i] = y[i] / H[i + ld * i];
            }
            for(int j = 0; j < k; j++)
                LAL_TRY(lal_axpy_f64(c, m, y[j], col(j), x));
            LAL_TRY(A.residual(c, x, b, res));
            LAL_TRY(P.apply(c, m, res, z));
            LAL_TRY(lal_dot_f64(c, m, z, z, &vv));
            rn = std::sqrt(vv);
            if(lal::converges(rn, init, o.rel_tol, o.abs_tol))
            {
                conv = true;
                break;
            }
            LAL_TRY(div_scalar(c, m, z, rn, col(0)));
            std::fill(y.begin(), y.end(), 0.0);
            y[0] = rn;
            iter++;
        }
    }
    if(info)
    {
        info->iterations       = iter;
        info->outer_iterations = steps;
        info->converged        = conv;
        info->initial_res_norm = init;
        info->final_res_norm   = rn;
    }
    return LAL_OK;
}
