// Test-infrastructure driver around LinAlgLib's own host library.
//
// oracle/Makefile compiles LinAlgLib's host sources where they lie
// (library/src/**/*.cpp, tracing off) together with this file into
// oracle/_ref/libref.so.  It exists to pin the C oracle (lal_oracle.c) and to
// serve as bench.py's reference CPU baseline.  No LinAlgLib source is copied
// here: this file only calls the library's public API.
//
// The SA-AMG preconditioner below is the composition the reference's public
// API allows (there is no AMG preconditioner class in LinAlgLib): zero x, then
// amg_solve(h, x, rhs, n1, n2, vcycle, jacobi, {max_cycle = 0}), which runs
// exactly one V-cycle (library/src/iterative_solvers/amg/amg.cpp:516-571,
// iter_control.cpp:88-91).
#include "linalg.h"

#include <cstdio>
#include <cstring>
#include <fcntl.h>
#include <iostream>
#include <omp.h>
#include <unistd.h>
#include <vector>

namespace
{
    // LinAlgLib prints progress (setup levels, aggregates, timings) to stdout
    // and, without a device backend, a warning per freed vector to stdout.
    struct quiet
    {
        int saved1, saved2;
        quiet()
        {
            std::cout.flush();
            std::fflush(nullptr);
            saved1 = dup(1);
            saved2 = dup(2);
            int nd = open("/dev/null", O_WRONLY);
            dup2(nd, 1);
            dup2(nd, 2);
            close(nd);
        }
        ~quiet()
        {
            std::cout.flush();
            std::cerr.flush();
            std::fflush(nullptr);
            dup2(saved1, 1);
            dup2(saved2, 2);
            close(saved1);
            close(saved2);
        }
    };

    linalg::csr_matrix<double>* make_csr(int m, int n, int nnz, const int* rp, const int* ci,
                                         const double* v)
    {
        std::vector<int>    r(rp, rp + m + 1);
        std::vector<int>    c(ci, ci + nnz);
        std::vector<double> w(v, v + nnz);
        return new linalg::csr_matrix<double>(r, c, w, m, n, nnz);
    }

    struct ref_hier
    {
        linalg::hierarchy h;
    };

    class amg_precond : public linalg::preconditioner
    {
    public:
        const linalg::hierarchy* h;
        int                      n1, n2;
        void build(const linalg::csr_matrix<double>&) override {}
        void solve(const linalg::vector<double>& rhs, linalg::vector<double>& x) const override
        {
            x.zeros();
            linalg::iter_control c;
            c.max_cycle = 0;
            linalg::amg_solve(
                *h, x, rhs, n1, n2, linalg::cycle::vcycle, linalg::smoother::jacobi, c);
        }
        void move_to_device() override {}
        void move_to_host() override {}
        bool is_on_host() const override
        {
            return true;
        }
    };

    const linalg::csr_matrix<double>* pick(const ref_hier* r, int level, int kind)
    {
        if(level < 0)
            return nullptr;
        if(kind == 0)
            return level <= r->h.total_levels ? &r->h.A_cs[level] : nullptr;
        if(level >= r->h.total_levels)
            return nullptr;
        return kind == 1 ? &r->h.prolongations[level] : &r->h.restrictions[level];
    }
}

extern "C" {

void ref_set_threads(int n)
{
    omp_set_num_threads(n);
}

void* ref_saamg_setup(int m, int n, int nnz, const int* rp, const int* ci, const double* v,
                      int max_level)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    ref_hier*                   r = new ref_hier;
    linalg::saamg_setup(*A, max_level, r->h);
    delete A;
    return r;
}

void* ref_uaamg_setup(int m, int n, int nnz, const int* rp, const int* ci, const double* v,
                      int max_level)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    ref_hier*                   r = new ref_hier;
    linalg::uaamg_setup(*A, max_level, r->h);
    delete A;
    return r;
}

void ref_hierarchy_free(void* h)
{
    quiet q;
    delete static_cast<ref_hier*>(h);
}

int ref_hierarchy_levels(void* h)
{
    return static_cast<ref_hier*>(h)->h.total_levels;
}

int ref_hierarchy_dims(void* h, int level, int kind, int* m, int* n, int* nnz)
{
    quiet q;
    const linalg::csr_matrix<double>* A = pick(static_cast<ref_hier*>(h), level, kind);
    if(!A)
        return -1;
    *m   = A->get_m();
    *n   = A->get_n();
    *nnz = A->get_nnz();
    return 0;
}

void ref_hierarchy_copy(void* h, int level, int kind, int* rp, int* ci, double* v)
{
    quiet q;
    const linalg::csr_matrix<double>* A = pick(static_cast<ref_hier*>(h), level, kind);
    if(!A)
        return;
    std::memcpy(rp, A->get_row_ptr(), sizeof(int) * (A->get_m() + 1));
    std::memcpy(ci, A->get_col_ind(), sizeof(int) * A->get_nnz());
    std::memcpy(v, A->get_val(), sizeof(double) * A->get_nnz());
}

// One SA-AMG preconditioner application z = M^-1 r.
void ref_amg_precond(void* h, const double* r, double* z, int n1, int n2)
{
    quiet q;
    ref_hier*             rh = static_cast<ref_hier*>(h);
    int                   n  = rh->h.A_cs[0].get_m();
    linalg::vector<double> rv(std::vector<double>(r, r + n));
    linalg::vector<double> zv(n);
    amg_precond            p;
    p.h  = &rh->h;
    p.n1 = n1;
    p.n2 = n2;
    p.solve(rv, zv);
    std::memcpy(z, zv.get_vec(), sizeof(double) * n);
}

// cg_solver::solve with the composed SA-AMG preconditioner.  x is the initial
// guess on entry and the solution on exit.
int ref_pcg_saamg(void* h, double* x, const double* b, double rel_tol, double abs_tol,
                  int max_iter, int n1, int n2)
{
    quiet q;
    ref_hier*              rh = static_cast<ref_hier*>(h);
    const auto&            A  = rh->h.A_cs[0];
    int                    n  = A.get_m();
    linalg::vector<double> xv(std::vector<double>(x, x + n));
    linalg::vector<double> bv(std::vector<double>(b, b + n));
    amg_precond            p;
    p.h  = &rh->h;
    p.n1 = n1;
    p.n2 = n2;
    linalg::cg_solver cg;
    cg.build(A);
    linalg::iter_control c;
    c.rel_tol  = rel_tol;
    c.abs_tol  = abs_tol;
    c.max_iter = max_iter;
    int it     = cg.solve(A, xv, bv, &p, c);
    std::memcpy(x, xv.get_vec(), sizeof(double) * n);
    return it;
}

// cg_solver::solve(A, x, b, precond, control) with the Krylov operator A given
// separately from the hierarchy behind the composed SA-AMG preconditioner.
int ref_pcg_saamg_op(int m, int nnz, const int* rp, const int* ci, const double* v, void* h,
                     double* x, const double* b, double rel_tol, double abs_tol, int max_iter,
                     int n1, int n2)
{
    quiet q;
    ref_hier*                   rh = static_cast<ref_hier*>(h);
    linalg::csr_matrix<double>* A  = make_csr(m, m, nnz, rp, ci, v);
    linalg::vector<double>      xv(std::vector<double>(x, x + m));
    linalg::vector<double>      bv(std::vector<double>(b, b + m));
    amg_precond                 p;
    p.h  = &rh->h;
    p.n1 = n1;
    p.n2 = n2;
    linalg::cg_solver cg;
    cg.build(*A);
    linalg::iter_control c;
    c.rel_tol  = rel_tol;
    c.abs_tol  = abs_tol;
    c.max_iter = max_iter;
    int it     = cg.solve(*A, xv, bv, &p, c);
    std::memcpy(x, xv.get_vec(), sizeof(double) * m);
    delete A;
    return it;
}

// amg_solve with the given cycle (0 V, 1 W, 2 F) and smoother (0 Jacobi,
// 1 Gauss-Seidel, 2 symmetric Gauss-Seidel, 3 SOR, 4 SSOR); x in place.
int ref_amg_solve(void* h, double* x, const double* b, int n1, int n2, int cycle, int smoother,
                  double rel_tol, double abs_tol, int max_cycle)
{
    quiet q;
    ref_hier*              rh = static_cast<ref_hier*>(h);
    int                    n  = rh->h.A_cs[0].get_m();
    linalg::vector<double> xv(std::vector<double>(x, x + n));
    linalg::vector<double> bv(std::vector<double>(b, b + n));
    linalg::iter_control   c;
    c.rel_tol   = rel_tol;
    c.abs_tol   = abs_tol;
    c.max_cycle = max_cycle;
    int k       = linalg::amg_solve(rh->h, xv, bv, n1, n2, static_cast<linalg::cycle>(cycle),
                                    static_cast<linalg::smoother>(smoother), c);
    std::memcpy(x, xv.get_vec(), sizeof(double) * n);
    return k;
}

} // extern "C"

namespace
{
    // precond: 0 none, 1 jacobi_precond, 2 one SA-AMG V-cycle of hierarchy h
    template <typename Solver, typename Build>
    int run_krylov(Build build, int m, int nnz, const int* rp, const int* ci, const double* v,
                   double* x, const double* b, int precond, void* h, double rel_tol,
                   double abs_tol, int max_iter)
    {
        linalg::csr_matrix<double>* A = make_csr(m, m, nnz, rp, ci, v);
        linalg::vector<double>      xv(std::vector<double>(x, x + m));
        linalg::vector<double>      bv(std::vector<double>(b, b + m));
        Solver                      solver;
        build(solver, *A);
        linalg::iter_control c;
        c.rel_tol  = rel_tol;
        c.abs_tol  = abs_tol;
        c.max_iter = max_iter;
        linalg::jacobi_precond jp;
        amg_precond            ap;
        const linalg::preconditioner* p = nullptr;
        if(precond == 1)
        {
            jp.build(*A);
            p = &jp;
        }
        else if(precond == 2)
        {
            ap.h  = &static_cast<ref_hier*>(h)->h;
            ap.n1 = 1;
            ap.n2 = 1;
            p     = &ap;
        }
        int it = solver.solve(*A, xv, bv, p, c);
        std::memcpy(x, xv.get_vec(), sizeof(double) * m);
        delete A;
        return it;
    }
}

extern "C" {

// bicgstab_solver::solve (krylov/bicgstab.cpp:64-249); x in place.
int ref_bicgstab(int m, int nnz, const int* rp, const int* ci, const double* v, double* x,
                 const double* b, int precond, void* h, double rel_tol, double abs_tol,
                 int max_iter)
{
    quiet q;
    return run_krylov<linalg::bicgstab_solver>(
        [](linalg::bicgstab_solver& s, const linalg::csr_matrix<double>& A) { s.build(A); }, m,
        nnz, rp, ci, v, x, b, precond, h, rel_tol, abs_tol, max_iter);
}

// gmres_solver::solve (krylov/gmres.cpp:224-540) with the given restart; x in place.
int ref_gmres(int m, int nnz, const int* rp, const int* ci, const double* v, double* x,
              const double* b, int restart, int precond, void* h, double rel_tol,
              double abs_tol, int max_iter)
{
    quiet q;
    return run_krylov<linalg::gmres_solver>(
        [restart](linalg::gmres_solver& s, const linalg::csr_matrix<double>& A) {
            s.build(A, restart);
        },
        m, nnz, rp, ci, v, x, b, precond, h, rel_tol, abs_tol, max_iter);
}

// Unpreconditioned cg_solver::solve.
int ref_cg(int m, int nnz, const int* rp, const int* ci, const double* v, double* x,
           const double* b, double rel_tol, double abs_tol, int max_iter)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, m, nnz, rp, ci, v);
    linalg::vector<double>      xv(std::vector<double>(x, x + m));
    linalg::vector<double>      bv(std::vector<double>(b, b + m));
    linalg::cg_solver           cg;
    cg.build(*A);
    linalg::iter_control c;
    c.rel_tol  = rel_tol;
    c.abs_tol  = abs_tol;
    c.max_iter = max_iter;
    int it     = cg.solve(*A, xv, bv, nullptr, c);
    std::memcpy(x, xv.get_vec(), sizeof(double) * m);
    delete A;
    return it;
}

void ref_csrmv(int m, int n, int nnz, double alpha, const int* rp, const int* ci,
               const double* v, const double* x, double beta, double* y)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    linalg::vector<double>      xv(std::vector<double>(x, x + n));
    linalg::vector<double>      yv(std::vector<double>(y, y + m));
    linalg::csrmv_descr*        d = nullptr;
    linalg::create_csrmv_descr(&d);
    linalg::csrmv_analysis(*A, linalg::csrmv_algorithm::default_algorithm, d);
    linalg::csrmv_solve(alpha, *A, xv, beta, yv, linalg::csrmv_algorithm::default_algorithm, d);
    linalg::destroy_csrmv_descr(d);
    std::memcpy(y, yv.get_vec(), sizeof(double) * m);
    delete A;
}

void ref_residual(int m, int n, int nnz, const int* rp, const int* ci, const double* v,
                  const double* x, const double* b, double* r)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    linalg::vector<double>      xv(std::vector<double>(x, x + n));
    linalg::vector<double>      bv(std::vector<double>(b, b + m));
    linalg::vector<double>      rv(m);
    linalg::compute_residual(*A, xv, bv, rv);
    std::memcpy(r, rv.get_vec(), sizeof(double) * m);
    delete A;
}

double ref_dot(int n, const double* x, const double* y)
{
    quiet q;
    linalg::vector<double> xv(std::vector<double>(x, x + n));
    linalg::vector<double> yv(std::vector<double>(y, y + n));
    return linalg::dot_product(xv, yv);
}

void ref_strength(int m, int n, int nnz, const int* rp, const int* ci, const double* v,
                  double eps, int* conn)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    linalg::vector<int>         c(nnz);
    c.zeros();
    linalg::compute_strong_connections(*A, eps, c);
    std::memcpy(conn, c.get_vec(), sizeof(int) * nnz);
    delete A;
}

void ref_aggregates(int m, int n, int nnz, const int* rp, const int* ci, const double* v,
                    const int* conn, int64_t* aggregates)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    linalg::vector<int>         c(std::vector<int>(conn, conn + nnz));
    linalg::vector<int64_t>     agg(m);
    agg.zeros();
    linalg::vector<int64_t> roots;
    linalg::compute_aggregates_using_pmis(*A, c, agg, roots);
    std::memcpy(aggregates, agg.get_vec(), sizeof(int64_t) * m);
    delete A;
}

// C = A * B through csr_matrix::multiply_by_matrix; returns an opaque handle.
void* ref_csrgemm(int ma, int na, int nnza, const int* rpa, const int* cia, const double* va,
                  int mb, int nb, int nnzb, const int* rpb, const int* cib, const double* vb)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(ma, na, nnza, rpa, cia, va);
    linalg::csr_matrix<double>* B = make_csr(mb, nb, nnzb, rpb, cib, vb);
    auto*                       C = new linalg::csr_matrix<double>();
    A->multiply_by_matrix(*C, *B);
    delete A;
    delete B;
    return C;
}

void* ref_transpose(int m, int n, int nnz, const int* rp, const int* ci, const double* v)
{
    quiet q;
    linalg::csr_matrix<double>* A = make_csr(m, n, nnz, rp, ci, v);
    auto*                       T = new linalg::csr_matrix<double>();
    A->transpose(*T);
    delete A;
    return T;
}

// csr_matrix::read_mtx into a new host matrix (nullptr when it refuses the file)
void* ref_read_mtx(const char* filename)
{
    quiet q;
    auto* A = new linalg::csr_matrix<double>();
    if(!A->read_mtx(filename))
    {
        delete A;
        return nullptr;
    }
    return A;
}

// csr_matrix::write_mtx of the given CSR arrays
int ref_write_mtx(const char* filename, int m, int n, int nnz, const int* rp, const int* ci,
                  const double* v)
{
    quiet q;
    linalg::csr_matrix<double>* A  = make_csr(m, n, nnz, rp, ci, v);
    const bool                  ok = A->write_mtx(filename);
    delete A;
    return ok ? 1 : 0;
}

void ref_csr_dims(void* p, int* m, int* n, int* nnz)
{
    quiet q;
    auto* A = static_cast<linalg::csr_matrix<double>*>(p);
    *m      = A->get_m();
    *n      = A->get_n();
    *nnz    = A->get_nnz();
}

void ref_csr_copy(void* p, int* rp, int* ci, double* v)
{
    quiet q;
    auto* A = static_cast<linalg::csr_matrix<double>*>(p);
    std::memcpy(rp, A->get_row_ptr(), sizeof(int) * (A->get_m() + 1));
    std::memcpy(ci, A->get_col_ind(), sizeof(int) * A->get_nnz());
    std::memcpy(v, A->get_val(), sizeof(double) * A->get_nnz());
}

void ref_csr_free(void* p)
{
    quiet q;
    delete static_cast<linalg::csr_matrix<double>*>(p);
}

} // extern "C"
