// lal_linalg.hpp -- LinAlgLib's host C++ API for the SA-AMG-PCG path, on the
// MI355X C ABI (lal.h).  Header only; links against liblal.so, no HIP headers.
//
// The names and argument meaning follow LinAlgLib (paths relative to its
// library/ tree) so host code written against the reference reads the same:
//   csr_matrix<T>        include/csr_matrix.h:48-449   (device resident here)
//   vector<T>            include/vector.h:57-210       (device resident here)
//   iter_control         include/iterative_solvers/iter_control.h:43-108
//   hierarchy            include/iterative_solvers/amg/amg.h:65-144
//   saamg_setup          include/iterative_solvers/amg/saamg.h:195-196
//   amg_solve            include/iterative_solvers/amg/amg.h (src amg.cpp:516-571)
//   preconditioner       include/iterative_solvers/preconditioner/preconditioner.h:59-99
//   cg_solver            include/iterative_solvers/krylov/cg.h (src cg.cpp:45-241)
// saamg_precond is new: the reference has no AMG preconditioner class; it is
// the composition x.zeros(); amg_solve(h, x, rhs, 1, 1, vcycle, jacobi,
// {max_cycle = 0}) executed on the device.
//
// Errors keep the reference's convention -- print and carry on -- except
// that a failed device call also throws std::runtime_error, since there is no
// host path to fall back to.
#pragma once

#include "lal.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <memory>
#include <stdexcept>
#include <string>
#include <string>
#include <utility>
#include <vector>

namespace lal_linalg
{
    inline lal_handle context()
    {
        static lal_handle h = [] {
            lal_handle t = nullptr;
            int        s = lal_create(&t);
            if(s != LAL_OK)
                throw std::runtime_error(std::string("lal_create: ") + lal_status_string(s));
            return t;
        }();
        return h;
    }

    inline void check(int s, const char* where)
    {
        if(s != LAL_OK)
        {
            std::cout << "Error: " << where << ": " << lal_status_string(s) << std::endl;
            throw std::runtime_error(std::string(where) + ": " + lal_status_string(s));
        }
    }

    inline void synchronize()
    {
        check(lal_sync(context()), "synchronize");
    }

    // ------------------------------------------------------------------ vector
    template <typename T>
    class vector
    {
        T*     d    = nullptr;
        size_t size = 0;

    public:
        vector() = default;
        explicit vector(size_t n)
        {
            resize(n);
        }
        vector(size_t n, T val)
        {
            resize(n);
            fill(val);
        }
        explicit vector(const std::vector<T>& v)
        {
            resize(v.size());
            check(lal_memcpy_h2d(context(), d, v.data(), sizeof(T) * size), "vector(h2d)");
        }
        vector(const vector&)            = delete;
        vector& operator=(const vector&) = delete;
        vector(vector&& o) noexcept : d(o.d), size(o.size)
        {
            o.d    = nullptr;
            o.size = 0;
        }
        ~vector()
        {
            lal_free(d);
        }

        bool is_on_host() const
        {
            return false;
        }
        size_t get_size() const
        {
            return size;
        }
        T* get_vec()
        {
            return d;
        }
        const T* get_vec() const
        {
            return d;
        }
        void resize(size_t n)
        {
            if(n == size)
                return;
            lal_free(d);
            d    = nullptr;
            size = n;
            check(lal_alloc((void**)&d, sizeof(T) * n), "vector::resize");
        }
        void resize(size_t n, T val)
        {
            resize(n);
            fill(val);
        }
        void clear()
        {
            lal_free(d);
            d    = nullptr;
            size = 0;
        }
        void fill(T val);
        void zeros()
        {
            fill(T(0));
        }
        void ones()
        {
            fill(T(1));
        }
        // vector<T>::rand (vector.cpp:182-199, host_memory.cpp:74-85): the same
        // formula over the C library's rand(), filled on the host and uploaded
        void rand(T min, T max)
        {
            std::vector<T> h(size);
            for(size_t i = 0; i < size; i++)
                h[i] = min + static_cast<T>(std::rand()) / (static_cast<T>(RAND_MAX / (max - min)));
            check(lal_memcpy_h2d(context(), d, h.data(), sizeof(T) * size), "vector::rand");
        }
        // The data lives on the device for its whole life: residency moves are
        // no-ops (vector.cpp's move_to_device / move_to_host).
        void move_to_device() {}
        void move_to_host() {}
        void copy_from(const vector& x)
        {
            resize(x.get_size());
            check(lal_memcpy_d2d(context(), d, x.d, sizeof(T) * size), "vector::copy_from");
        }
        std::vector<T> to_host() const
        {
            std::vector<T> h(size);
            check(lal_memcpy_d2h(context(), h.data(), d, sizeof(T) * size), "vector::to_host");
            return h;
        }
    };

    template <>
    inline void vector<double>::fill(double v)
    {
        check(lal_fill_f64(context(), d, size, v), "vector::fill");
    }
    template <>
    inline void vector<float>::fill(float v)
    {
        check(lal_fill_f32(context(), d, size, v), "vector::fill");
    }
    template <>
    inline void vector<int>::fill(int v)
    {
        check(lal_fill_i32(context(), d, size, v), "vector::fill");
    }
    template <>
    inline void vector<int64_t>::fill(int64_t v)
    {
        check(lal_fill_i64(context(), d, size, v), "vector::fill");
    }

    // ------------------------------------------------------------- math (BLAS-1)
    inline void axpy(double alpha, const vector<double>& x, vector<double>& y)
    {
        check(lal_axpy_f64(context(), (int)x.get_size(), alpha, x.get_vec(), y.get_vec()), "axpy");
    }
    inline void axpby(double alpha, const vector<double>& x, double beta, vector<double>& y)
    {
        check(lal_axpby_f64(context(), (int)x.get_size(), alpha, x.get_vec(), beta, y.get_vec()),
              "axpby");
    }
    inline double dot_product(const vector<double>& x, const vector<double>& y)
    {
        double r = 0;
        check(lal_dot_f64(context(), (int)x.get_size(), x.get_vec(), y.get_vec(), &r),
              "dot_product");
        return r;
    }
    inline double norm_inf(const vector<double>& x)
    {
        double r = 0;
        check(lal_norm_inf_f64(context(), (int)x.get_size(), x.get_vec(), &r), "norm_inf");
        return r;
    }
    template <typename T>
    inline void exclusive_scan(vector<T>& x);
    template <>
    inline void exclusive_scan<int>(vector<int>& x)
    {
        check(lal_exclusive_scan_i32(context(), (int)x.get_size(), x.get_vec()), "exclusive_scan");
    }
    template <>
    inline void exclusive_scan<int64_t>(vector<int64_t>& x)
    {
        check(lal_exclusive_scan_i64(context(), (int)x.get_size(), x.get_vec()), "exclusive_scan");
    }

    // -------------------------------------------------------------- csr_matrix
    // Matrix Market coordinate files on the host (csr_matrix.cpp:523-794):
    // "matrix coordinate real|integer general|symmetric" is read, 1-based
    // entries become a CSR matrix with columns sorted inside rows, a symmetric
    // file's (i, j), i != j, also yields (j, i), and repeated (i, j) entries
    // are summed (in file order).  Other objects, formats, fields and
    // symmetries are refused with a message on stderr, as in the reference.
    namespace mtx
    {
        inline bool read(const std::string& filename, int& m, int& n, std::vector<int>& row_ptr,
                         std::vector<int>& col_ind, std::vector<double>& val)
        {
            std::ifstream in(filename);
            if(!in)
            {
                std::cout << "Error: Could not read file contents" << std::endl;
                return false;
            }
            std::string line, tok;
            std::getline(in, line);
            std::istringstream h(line);
            std::string        banner, object, format, field, symmetry;
            h >> banner >> object >> format >> field >> symmetry;
            if(object != "matrix")
            {
                std::cerr << "Error: Not a matrix market file." << std::endl;
                return false;
            }
            if(format != "coordinate")
            {
                std::cerr << "Error: Only 'coordinate' format is supported (not 'array')."
                          << std::endl;
                return false;
            }
            if(field != "real" && field != "integer")
            {
                std::cerr << "Error: unsupported data type '" << field << "'." << std::endl;
                return false;
            }
            if(symmetry != "general" && symmetry != "symmetric")
            {
                std::cerr << "Error: unsupported symmetry '" << symmetry << "'." << std::endl;
                return false;
            }
            const bool sym = symmetry == "symmetric";
            while(std::getline(in, line) && !line.empty() && line[0] == '%')
            {
            }
            int64_t entries = 0;
            {
                std::istringstream d(line);
                if(!(d >> m >> n >> entries))
                {
                    std::cerr << "Error: Failed to read matrix dimensions or number of non-zeros."
                              << std::endl;
                    return false;
                }
            }
            struct entry
            {
                int    r, c;
                double v;
            };
            std::vector<entry> t;
            t.reserve(sym ? 2 * entries : entries);
            for(int64_t k = 0; k < entries; k++)
            {
                if(!std::getline(in, line))
                {
                    std::cerr << "Error: Unexpected end of file while reading elements (expected "
                              << entries << " elements)." << std::endl;
                    return false;
                }
                std::istringstream e(line);
                int    r, c;
                double v;
                if(!(e >> r >> c >> v) || r < 1 || c < 1 || r > m || c > n)
                {
                    std::cerr << "Error: malformed entry '" << line << "'." << std::endl;
                    return false;
                }
                t.push_back({r - 1, c - 1, v});
                if(sym && r != c)
                    t.push_back({c - 1, r - 1, v});
            }
            std::stable_sort(t.begin(), t.end(), [](const entry& a, const entry& b) {
                return a.r != b.r ? a.r < b.r : a.c < b.c;
            });
            row_ptr.assign(m + 1, 0);
            col_ind.clear();
            val.clear();
            for(size_t k = 0; k < t.size(); k++)
            {
                if(!col_ind.empty() && k > 0 && t[k].r == t[k - 1].r && t[k].c == t[k - 1].c)
                {
                    val.back() += t[k].v;
                    continue;
                }
                row_ptr[t[k].r + 1]++;
                col_ind.push_back(t[k].c);
                val.push_back(t[k].v);
            }
            for(int i = 0; i < m; i++)
                row_ptr[i + 1] += row_ptr[i];
            return true;
        }

        // Writes "real general" with 10 digits after the point, as the
        // reference's writer does (csr_matrix.cpp:736-794).
        inline bool write(const std::string& filename, int m, int n,
                          const std::vector<int>& row_ptr, const std::vector<int>& col_ind,
                          const std::vector<double>& val)
        {
            std::ofstream out(filename);
            if(!out.is_open())
            {
                std::cerr << "Error: Could not open file " << filename << " for writing."
                          << std::endl;
                return false;
            }
            const int nnz = m > 0 ? row_ptr[m] : 0;
            out << "%%MatrixMarket matrix coordinate real general\n";
            out << "% Created by C++ CSR to Matrix Market writer\n";
            out << m << " " << n << " " << nnz << "\n";
            out << std::fixed << std::setprecision(10);
            for(int i = 0; i < m; i++)
                for(int j = row_ptr[i]; j < row_ptr[i + 1]; j++)
                    out << (i + 1) << " " << (col_ind[j] + 1) << " " << val[j] << "\n";
            return true;
        }
    }

    template <typename T>
    class csr_matrix;

    template <>
    class csr_matrix<double>
    {
        int             m = 0, n = 0, nnz = 0;
        vector<int>     rp, ci;
        vector<double>  v;
        mutable lal_csrmv_descr descr = nullptr;
        int             exact = 0;

        lal_csrmv_descr analysed() const
        {
            if(!descr)
                check(lal_csrmv_analysis(context(), m, n, nnz, rp.get_vec(), ci.get_vec(),
                                         exact, &descr),
                      "csrmv_analysis");
            return descr;
        }

    public:
        // Backend extension (no LinAlgLib counterpart): with exact = true every
        // product and residual with this matrix sums rows sequentially in CSR
        // order, bit-identical to the host path (host_matrix_vector.cpp:35-84).
        void set_exact(bool e)
        {
            if((int)e != exact)
            {
                lal_csrmv_descr_destroy(descr);
                descr = nullptr;
                exact = e;
            }
        }
        // The SpMV analysis of this pattern (built on first use).
        lal_csrmv_descr descriptor() const
        {
            return analysed();
        }
        csr_matrix() = default;
        csr_matrix(const std::vector<int>& row_ptr, const std::vector<int>& col_ind,
                   const std::vector<double>& val, int m_, int n_, int nnz_)
            : m(m_), n(n_), nnz(nnz_), rp(row_ptr), ci(col_ind), v(val)
        {
        }
        csr_matrix(const csr_matrix&)            = delete;
        csr_matrix& operator=(const csr_matrix&) = delete;
        ~csr_matrix()
        {
            lal_csrmv_descr_destroy(descr);
        }

        bool is_on_host() const
        {
            return false;
        }
        int get_m() const
        {
            return m;
        }
        int get_n() const
        {
            return n;
        }
        int get_nnz() const
        {
            return nnz;
        }
        const int* get_row_ptr() const
        {
            return rp.get_vec();
        }
        const int* get_col_ind() const
        {
            return ci.get_vec();
        }
        const double* get_val() const
        {
            return v.get_vec();
        }
        int* get_row_ptr()
        {
            return rp.get_vec();
        }
        int* get_col_ind()
        {
            return ci.get_vec();
        }
        double* get_val()
        {
            return v.get_vec();
        }

        void move_to_device() {}
        void move_to_host() {}
        // csr_matrix<T>::copy_from (csr_matrix.cpp): pattern and values
        void copy_from(const csr_matrix& A)
        {
            resize(A.m, A.n, A.nnz);
            rp.copy_from(A.rp);
            ci.copy_from(A.ci);
            v.copy_from(A.v);
        }

        void resize(int m_, int n_, int nnz_)
        {
            lal_csrmv_descr_destroy(descr);
            descr = nullptr;
            m     = m_;
            n     = n_;
            nnz   = nnz_;
            rp.resize(m + 1);
            ci.resize(nnz);
            v.resize(nnz);
        }

        // y = A x  (csr_matrix.cpp:302-315)
        void multiply_by_vector(vector<double>& y, const vector<double>& x) const
        {
            check(lal_csrmv_f64(context(), analysed(), m, n, nnz, 1.0, get_row_ptr(),
                                get_col_ind(), get_val(), x.get_vec(), 0.0, y.get_vec()),
                  "multiply_by_vector");
        }
        // y = A x + y  (csr_matrix.cpp:317-330)
        void multiply_by_vector_and_add(vector<double>& y, const vector<double>& x) const
        {
            check(lal_csrmv_f64(context(), analysed(), m, n, nnz, 1.0, get_row_ptr(),
                                get_col_ind(), get_val(), x.get_vec(), 1.0, y.get_vec()),
                  "multiply_by_vector_and_add");
        }
        void extract_diagonal(vector<double>& diag) const
        {
            check(lal_diagonal_f64(context(), m, get_row_ptr(), get_col_ind(), get_val(),
                                   diag.get_vec()),
                  "extract_diagonal");
        }
        // C = this * B  (csr_matrix.cpp:332-353)
        void multiply_by_matrix(csr_matrix& C, const csr_matrix& B) const
        {
            C.resize(m, B.n, 0);
            int nnz_c = 0;
            check(lal_csrgemm_nnz(context(), m, n, B.n, nnz, get_row_ptr(), get_col_ind(), B.nnz,
                                  B.get_row_ptr(), B.get_col_ind(), C.get_row_ptr(), &nnz_c),
                  "csrgemm_nnz");
            C.nnz = nnz_c;
            C.ci.resize(nnz_c);
            C.v.resize(nnz_c);
            check(lal_csrgemm_solve_f64(context(), m, n, B.n, nnz, get_row_ptr(), get_col_ind(),
                                        get_val(), B.nnz, B.get_row_ptr(), B.get_col_ind(),
                                        B.get_val(), C.get_row_ptr(), C.get_col_ind(),
                                        C.get_val()),
                  "csrgemm_solve");
        }
        // csr_matrix::read_mtx / write_mtx (csr_matrix.cpp:523-794): parsed on
        // the host, then the arrays move to the device.
        bool read_mtx(const std::string& filename)
        {
            int                 m_ = 0, n_ = 0;
            std::vector<int>    hrp, hci;
            std::vector<double> hv;
            if(!mtx::read(filename, m_, n_, hrp, hci, hv))
                return false;
            resize(m_, n_, (int)hci.size());
            if(!hci.empty())
            {
                check(lal_memcpy_h2d(context(), ci.get_vec(), hci.data(), sizeof(int) * hci.size()),
                      "read_mtx");
                check(lal_memcpy_h2d(context(), v.get_vec(), hv.data(), sizeof(double) * hv.size()),
                      "read_mtx");
            }
            check(lal_memcpy_h2d(context(), rp.get_vec(), hrp.data(), sizeof(int) * hrp.size()),
                  "read_mtx");
            return true;
        }
        bool write_mtx(const std::string& filename) const
        {
            return mtx::write(filename, m, n, rp.to_host(), nnz ? ci.to_host() : std::vector<int>(),
                              nnz ? v.to_host() : std::vector<double>());
        }

        // T = this^T  (csr_matrix.cpp:437-447)
        void transpose(csr_matrix& Tr) const
        {
            Tr.resize(n, m, nnz);
            check(lal_transpose_f64(context(), m, n, nnz, get_row_ptr(), get_col_ind(), get_val(),
                                    Tr.get_row_ptr(), Tr.get_col_ind(), Tr.get_val()),
                  "transpose");
        }
    };

    // r = b - A x  (linalg_math.cpp:110-119)
    inline void compute_residual(const csr_matrix<double>& A, const vector<double>& x,
                                 const vector<double>& b, vector<double>& r)
    {
        check(lal_residual_f64(context(), nullptr, A.get_m(), A.get_n(), A.get_nnz(),
                               A.get_row_ptr(), A.get_col_ind(), A.get_val(), x.get_vec(),
                               b.get_vec(), r.get_vec()),
              "compute_residual");
    }

    // ------------------------------------------------------------ iter_control
    struct iter_control
    {
        double rel_tol   = 1e-06;
        double abs_tol   = 1e-06;
        int    max_iter  = 1000;
        int    max_cycle = 200;

        // iterative_solvers/iter_control.cpp:38-78
        bool residual_converges(double residual_norm, double initial_residual_norm) const
        {
            if(std::isnan(residual_norm))
            {
                std::cout << "Error: NaN detected in residual norm" << std::endl;
                return false;
            }
            if(std::isnan(initial_residual_norm))
            {
                std::cout << "Error: NaN detected in initial residual norm" << std::endl;
                return false;
            }
            if(residual_norm <= abs_tol)
                return true;
            return residual_norm / initial_residual_norm <= rel_tol;
        }
        bool exceed_max_iter(int iter) const
        {
            return iter > max_iter;
        }
        bool exceed_max_cycle(int cycle) const
        {
            return cycle > max_cycle;
        }
    };

    // --------------------------------------------------------------------- AMG
    // include/iterative_solvers/amg/amg.h:155-243
    enum class cycle
    {
        vcycle,
        wcycle,
        fcycle
    };
    enum class smoother
    {
        jacobi,
        gauss_seidel,
        symmetric_gauss_seidel,
        SOR,
        SSOR
    };

    struct hierarchy
    {
        lal_hierarchy h            = nullptr;
        int           total_levels = 0;
        hierarchy()                = default;
        hierarchy(const hierarchy&) = delete;
        hierarchy& operator=(const hierarchy&) = delete;
        ~hierarchy()
        {
            lal_hierarchy_destroy(h);
        }
        bool is_on_host() const
        {
            return false;
        }
    };

    // saamg.cpp:210-282 on the device; eps = 0.08, relax = 2/3 as the reference.
    inline void saamg_setup(const csr_matrix<double>& A, int max_level, hierarchy& h,
                            bool build_f32 = false)
    {
        lal_hierarchy_destroy(h.h);
        h.h = nullptr;
        lal_saamg_options o;
        lal_saamg_default_options(&o);
        o.max_levels = max_level < 20 ? max_level : 20;
        o.build_f32  = build_f32 ? 1 : 0;
        check(lal_saamg_setup(context(), A.get_m(), A.get_n(), A.get_nnz(), A.get_row_ptr(),
                              A.get_col_ind(), A.get_val(), &o, &h.h),
              "saamg_setup");
        check(lal_hierarchy_levels(h.h, &h.total_levels), "saamg_setup");
    }

    // uaamg.cpp:125-224 on the device (eps from 0.001, one 1.0 per row of P).
    inline void uaamg_setup(const csr_matrix<double>& A, int max_level, hierarchy& h)
    {
        lal_hierarchy_destroy(h.h);
        h.h = nullptr;
        lal_saamg_options o;
        lal_uaamg_default_options(&o);
        o.max_levels = max_level < 20 ? max_level : 20;
        check(lal_uaamg_setup(context(), A.get_m(), A.get_n(), A.get_nnz(), A.get_row_ptr(),
                              A.get_col_ind(), A.get_val(), &o, &h.h),
              "uaamg_setup");
        check(lal_hierarchy_levels(h.h, &h.total_levels), "uaamg_setup");
    }

    // amg.cpp:516-571: cycles on x in place until the inf-norm residual
    // converges or more than control.max_cycle cycles ran; returns the cycle
    // count as the reference does.
    inline int amg_solve(const hierarchy& h, vector<double>& x, const vector<double>& b, int n1,
                         int n2, cycle cycle_type, smoother smoother_type,
                         iter_control control)
    {
        lal_solve_options o;
        lal_solve_default_options(&o);
        o.rel_tol  = control.rel_tol;
        o.abs_tol  = control.abs_tol;
        o.max_iter = control.max_cycle;
        lal_solve_info info{};
        check(lal_amg_solve(context(), h.h, b.get_vec(), x.get_vec(), n1, n2,
                            static_cast<int>(cycle_type), static_cast<int>(smoother_type), &o,
                            &info),
              "amg_solve");
        return info.iterations;
    }

    // ---------------------------------------------------------- preconditioner
    class preconditioner
    {
    public:
        virtual ~preconditioner() {}
        virtual void build(const csr_matrix<double>& A)                        = 0;
        virtual void solve(const vector<double>& rhs, vector<double>& x) const = 0;
        virtual void move_to_device() {}
        virtual void move_to_host() {}
        virtual bool is_on_host() const
        {
            return false;
        }
    };

    // One SA-AMG V-cycle (1 pre + 1 post undamped Jacobi sweep) from x = 0.
    class saamg_precond : public preconditioner
    {
    public:
        hierarchy h;
        int       n1 = 1, n2 = 1;
        void build(const csr_matrix<double>& A) override
        {
            saamg_setup(A, 20, h);
        }
        void solve(const vector<double>& rhs, vector<double>& x) const override
        {
            check(lal_amg_precond_f64(context(), h.h, rhs.get_vec(), x.get_vec(), n1, n2),
                  "saamg_precond::solve");
        }
    };

    // z = r / a_ii with a_ii the first diagonal entry of each row
    // (preconditioner.cpp:41-58, host_math.cpp:53-64).
    class jacobi_precond : public preconditioner
    {
    public:
        vector<double> diag;
        void build(const csr_matrix<double>& A) override
        {
            diag.resize(A.get_m());
            diag.zeros();
            A.extract_diagonal(diag);
        }
        void solve(const vector<double>& rhs, vector<double>& x) const override
        {
            check(lal_jacobi_solve_f64(context(), (int)rhs.get_size(), rhs.get_vec(),
                                       diag.get_vec(), x.get_vec()),
                  "jacobi_precond::solve");
        }
    };

    namespace detail
    {
        // The device solvers' preconditioner kind for a preconditioner object.
        inline int precond_kind(const preconditioner* p, lal_hierarchy* h, const char* where)
        {
            *h = nullptr;
            if(!p)
                return LAL_PRECOND_NONE;
            if(dynamic_cast<const jacobi_precond*>(p))
                return LAL_PRECOND_JACOBI;
            if(auto* sa = dynamic_cast<const saamg_precond*>(p))
            {
                *h = sa->h.h;
                return LAL_PRECOND_SAAMG;
            }
            throw std::runtime_error(std::string(where) +
                                     ": this backend runs no, Jacobi or SA-AMG preconditioning");
        }

        inline lal_solve_options options(const iter_control& control, const preconditioner* p)
        {
            lal_solve_options o;
            lal_solve_default_options(&o);
            o.rel_tol  = control.rel_tol;
            o.abs_tol  = control.abs_tol;
            o.max_iter = control.max_iter;
            if(auto* sa = dynamic_cast<const saamg_precond*>(p))
            {
                o.n1 = sa->n1;
                o.n2 = sa->n2;
            }
            return o;
        }
    }

    // ---------------------------------------------------------- bicgstab_solver
    // krylov/bicgstab.cpp:64-231 on the device; returns the iteration count.
    class bicgstab_solver
    {
    public:
        void build(const csr_matrix<double>&) {}

        int solve_nonprecond(const csr_matrix<double>& A, vector<double>& x,
                             const vector<double>& b, iter_control control)
        {
            return solve_precond(A, x, b, nullptr, control);
        }
        int solve_precond(const csr_matrix<double>& A, vector<double>& x, const vector<double>& b,
                          const preconditioner* precond, iter_control control)
        {
            lal_hierarchy     h    = nullptr;
            const int         kind = detail::precond_kind(precond, &h, "bicgstab_solver");
            lal_solve_options o    = detail::options(control, precond);
            lal_solve_info    info{};
            check(lal_bicgstab_solve(context(), A.descriptor(), A.get_m(), A.get_nnz(),
                                     A.get_row_ptr(), A.get_col_ind(), A.get_val(), b.get_vec(),
                                     x.get_vec(), kind, h, &o, &info),
                  "bicgstab_solver::solve");
            return info.iterations;
        }
        int solve(const csr_matrix<double>& A, vector<double>& x, const vector<double>& b,
                  const preconditioner* precond, iter_control control)
        {
            return precond ? solve_precond(A, x, b, precond, control)
                           : solve_nonprecond(A, x, b, control);
        }
    };

    // ------------------------------------------------------------- gmres_solver
    // Restarted, left-preconditioned GMRES (krylov/gmres.cpp:46-530 as
    // intended); returns the number of restart cycles, as gmres.cpp does.
    class gmres_solver
    {
        int restart = 30;

    public:
        void build(const csr_matrix<double>&, int restart_)
        {
            restart = restart_;
        }

        int solve_nonprecond(const csr_matrix<double>& A, vector<double>& x,
                             const vector<double>& b, iter_control control)
        {
            return solve_precond(A, x, b, nullptr, control);
        }
        int solve_precond(const csr_matrix<double>& A, vector<double>& x, const vector<double>& b,
                          const preconditioner* precond, iter_control control)
        {
            lal_hierarchy     h    = nullptr;
            const int         kind = detail::precond_kind(precond, &h, "gmres_solver");
            lal_solve_options o    = detail::options(control, precond);
            lal_solve_info    info{};
            check(lal_gmres_solve(context(), A.descriptor(), A.get_m(), A.get_nnz(),
                                  A.get_row_ptr(), A.get_col_ind(), A.get_val(), b.get_vec(),
                                  x.get_vec(), restart, kind, h, &o, &info),
                  "gmres_solver::solve");
            return info.iterations;
        }
        int solve(const csr_matrix<double>& A, vector<double>& x, const vector<double>& b,
                  const preconditioner* precond, iter_control control)
        {
            return precond ? solve_precond(A, x, b, precond, control)
                           : solve_nonprecond(A, x, b, control);
        }
    };

    // --------------------------------------------------------------- cg_solver
    class cg_solver
    {
        vector<double> z, p, res;
        int            restart_iter = -1; // cg.cpp:45-49

    public:
        void build(const csr_matrix<double>& A)
        {
            p.resize(A.get_m());
            z.resize(A.get_m());
            res.resize(A.get_m());
        }

        // cg.cpp:136-223, every step a device call.  With an saamg_precond
        // the whole loop runs in the library instead (lal_pcg_solve_op_f64:
        // the Krylov products on A, the V-cycle of the preconditioner's
        // hierarchy, which may have been built for another matrix).
        int solve_precond(const csr_matrix<double>& A, vector<double>& x,
                          const vector<double>& b, const preconditioner* precond,
                          iter_control control)
        {
            if(auto* sa = dynamic_cast<const saamg_precond*>(precond))
            {
                lal_solve_options o;
                lal_solve_default_options(&o);
                o.rel_tol  = control.rel_tol;
                o.abs_tol  = control.abs_tol;
                o.max_iter = control.max_iter;
                o.n1       = sa->n1;
                o.n2       = sa->n2;
                lal_solve_info info{};
                check(lal_pcg_solve_op_f64(context(), A.descriptor(), A.get_m(), A.get_nnz(),
                                           A.get_row_ptr(), A.get_col_ind(), A.get_val(),
                                           sa->h.h, b.get_vec(), x.get_vec(), &o, &info),
                      "cg_solver::solve_precond");
                return info.iterations;
            }
            compute_residual(A, x, b, res);
            double initial_res_norm = norm_inf(res);
            precond->solve(res, z);
            p.copy_from(z);
            double gamma = dot_product(z, res);
            int    iter  = 0;
            while(!control.exceed_max_iter(iter))
            {
                A.multiply_by_vector(z, p);
                double alpha = gamma / dot_product(z, p);
                axpy(alpha, p, x);
                axpy(-1.0 * alpha, z, res);
                double res_norm = norm_inf(res);
                if(control.residual_converges(res_norm, initial_res_norm))
                    break;
                precond->solve(res, z);
                double old_gamma = gamma;
                gamma            = dot_product(z, res);
                double beta      = gamma / old_gamma;
                axpby(1.0, z, beta, p);
                iter++;
            }
            return iter;
        }

        // cg.cpp:60-134, every step a device call.  As in the reference, the
        // restart block there is not guarded by restart_iter != -1, and
        // iter % -1 == 0, so with the default restart_iter = -1 every
        // iteration after the first restarts from the true residual (the
        // method is then steepest descent); kept so that iteration counts match.
        int solve_nonprecond(const csr_matrix<double>& A, vector<double>& x,
                             const vector<double>& b, iter_control control)
        {
            compute_residual(A, x, b, res);
            const double initial_res_norm = norm_inf(res);
            p.copy_from(res);
            double gamma = dot_product(res, res);
            int    iter  = 0;
            while(!control.exceed_max_iter(iter))
            {
                if(iter > 0 && iter % restart_iter == 0)
                {
                    compute_residual(A, x, b, res);
                    p.copy_from(res);
                    gamma = dot_product(res, res);
                }
                A.multiply_by_vector(z, p);
                double alpha = gamma / dot_product(z, p);
                axpy(alpha, p, x);
                axpy(-1.0 * alpha, z, res);
                double res_norm = norm_inf(res);
                if(control.residual_converges(res_norm, initial_res_norm))
                    break;
                double old_gamma = gamma;
                gamma            = dot_product(res, res);
                double beta      = gamma / old_gamma;
                axpby(1.0, res, beta, p);
                iter++;
            }
            return iter;
        }

        int solve(const csr_matrix<double>& A, vector<double>& x, const vector<double>& b,
                  const preconditioner* precond, iter_control control)
        {
            if(!precond)
                return solve_nonprecond(A, x, b, control);
            return solve_precond(A, x, b, precond, control);
        }
    };
}
