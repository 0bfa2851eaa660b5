// The bodies of LinAlgLib's device shims (library/src/backend/device/*.h)
// over the MI355X library's C ABI (include/lal.h), one call each where a
// counterpart exists.  Compiled together with LinAlgLib's own sources in
// place of backend/device/device_*.cpp (device_vector.cpp stays the
// reference's), this is the drop-in a maintainer would make: the reference's
// front end -- backend_dispatch (utility.h:163-193), csr_matrix, vector, the
// Krylov solvers -- runs unchanged and reaches the GPU through these bodies.
//
// Shims outside the SA-AMG-PCG path (ILU/IC, triangular solves, SpGEAM, Ruiz
// scaling, SSOR fills, tridiagonal solvers, classical strength, CF
// splitting, min/max, double scan) print the message the reference prints
// when it has no device backend, as its own shims do.
//
// Test infrastructure only: nothing under linalglib_amd/ links this file.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <map>
#include <mutex>
#include <type_traits>

#include "lal.h"

#include "backend/device/device_amg_aggregation.h"
#include "backend/device/device_amg_strength.h"
#include "backend/device/device_axpy.h"
#include "backend/device/device_csr2csc.h"
#include "backend/device/device_csrgeam.h"
#include "backend/device/device_csrgemm.h"
#include "backend/device/device_csric0.h"
#include "backend/device/device_csrilu0.h"
#include "backend/device/device_csrtrsv.h"
#include "backend/device/device_extract.h"
#include "backend/device/device_math.h"
#include "backend/device/device_matrix_vector.h"
#include "backend/device/device_memory.h"
#include "backend/device/device_primitives.h"
#include "backend/device/device_ruiz_scaling.h"
#include "backend/device/device_scale.h"
#include "backend/device/device_ssor.h"
#include "backend/device/device_tridiagonal.h"

namespace
{
    // One context per process, created on first use (lal_create: the current
    // device, its own stream of default priority).
    lal_handle ctx()
    {
        static lal_handle h = [] {
            lal_handle t = nullptr;
            if(lal_create(&t) != LAL_OK)
                std::cout << "Error: lal_create: " << lal_last_error() << std::endl;
            return t;
        }();
        return h;
    }

    // The reference prints device errors and carries on (utility.h:172-183).
    void check(int s, const char* where)
    {
        if(s != LAL_OK)
            std::cout << "Error: " << where << ": " << lal_status_string(s) << " ("
                      << lal_last_error() << ")" << std::endl;
    }

    void unsupported(const char* fn)
    {
        std::cout << "Error: Not device backend available for the function " << fn
                  << std::endl;
    }

    // csrmv_descr is an empty struct in the reference
    // (descriptors/csrmv_descr_internal.h); the analysis of each descriptor
    // is kept beside it.
    std::mutex                                           g_mu;
    std::map<const linalg::csrmv_descr*, lal_csrmv_descr> g_csrmv;

    lal_csrmv_descr analysis_of(const linalg::csrmv_descr* d)
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto                        it = g_csrmv.find(d);
        return it == g_csrmv.end() ? nullptr : it->second;
    }

    template <typename T>
    int fill(T* x, size_t n, T v);
    template <>
    int fill<double>(double* x, size_t n, double v)
    {
        return lal_fill_f64(ctx(), x, n, v);
    }
    template <>
    int fill<float>(float* x, size_t n, float v)
    {
        return lal_fill_f32(ctx(), x, n, v);
    }
    template <>
    int fill<int32_t>(int32_t* x, size_t n, int32_t v)
    {
        return lal_fill_i32(ctx(), x, n, v);
    }
    template <>
    int fill<uint32_t>(uint32_t* x, size_t n, uint32_t v)
    {
        return lal_fill_i32(ctx(), reinterpret_cast<int32_t*>(x), n, (int32_t)v);
    }
    template <>
    int fill<int64_t>(int64_t* x, size_t n, int64_t v)
    {
        return lal_fill_i64(ctx(), x, n, v);
    }
}

// The context every shim runs on (its stream orders all device work of the
// reference's front end); a program adding its own lal_* calls uses it too.
extern "C" lal_handle lal_dropin_context()
{
    return ctx();
}

// ---------------------------------------------------------------- memory
// backend/device/device_memory.h:34-52

template <typename T>
void linalg::device_allocate(T** ptr, size_t size)
{
    ctx();
    void* p = nullptr;
    check(lal_alloc(&p, size * sizeof(T)), "device_allocate");
    *ptr = static_cast<T*>(p);
}

template <typename T>
void linalg::device_free(T* ptr)
{
    check(lal_free(ptr), "device_free");
}

template <typename T>
void linalg::copy_h2d(T* dest, const T* src, size_t size)
{
    check(lal_memcpy_h2d(ctx(), dest, src, size * sizeof(T)), "copy_h2d");
}

template <typename T>
void linalg::copy_d2h(T* dest, const T* src, size_t size)
{
    check(lal_memcpy_d2h(ctx(), dest, src, size * sizeof(T)), "copy_d2h");
}

template <typename T>
void linalg::copy_d2d(T* dest, const T* src, size_t size)
{
    check(lal_memcpy_d2d(ctx(), dest, src, size * sizeof(T)), "copy_d2d");
}

template <typename T>
void linalg::device_fill(T* data, size_t size, T val)
{
    check(fill<T>(data, size, val), "device_fill");
}

void linalg::device_sync()
{
    check(lal_sync(ctx()), "device_sync");
}

#define LAL_SHIM_MEMORY(T)                                                    \
    template void linalg::device_allocate<T>(T * *ptr, size_t size);          \
    template void linalg::device_free<T>(T * ptr);                            \
    template void linalg::copy_h2d<T>(T * dest, const T* src, size_t size);   \
    template void linalg::copy_d2h<T>(T * dest, const T* src, size_t size);   \
    template void linalg::copy_d2d<T>(T * dest, const T* src, size_t size);   \
    template void linalg::device_fill<T>(T * data, size_t size, T val);
LAL_SHIM_MEMORY(uint32_t)
LAL_SHIM_MEMORY(int32_t)
LAL_SHIM_MEMORY(int64_t)
LAL_SHIM_MEMORY(float)
LAL_SHIM_MEMORY(double)
#undef LAL_SHIM_MEMORY

// ----------------------------------------------------------------- BLAS-1
// backend/device/device_axpy.h:33-41, device_math.h:54-60

void linalg::device_axpy(double alpha, const vector<double>& x, vector<double>& y)
{
    check(lal_axpy_f64(ctx(), (int)x.get_size(), alpha, x.get_vec(), y.get_vec()), "axpy");
}

void linalg::device_axpby(double alpha, const vector<double>& x, double beta, vector<double>& y)
{
    check(lal_axpby_f64(ctx(), (int)x.get_size(), alpha, x.get_vec(), beta, y.get_vec()),
          "axpby");
}

void linalg::device_axpbypgz(double                alpha,
                             const vector<double>& x,
                             double                beta,
                             const vector<double>& y,
                             double                gamma,
                             vector<double>&       z)
{
    check(lal_axpbypgz_f64(ctx(), (int)x.get_size(), alpha, x.get_vec(), beta, y.get_vec(),
                           gamma, z.get_vec()),
          "axpbypgz");
}

double linalg::device_dot_product(const vector<double>& x, const vector<double>& y)
{
    double r = 0.0;
    check(lal_dot_f64(ctx(), (int)x.get_size(), x.get_vec(), y.get_vec(), &r), "dot_product");
    return r;
}

double linalg::device_norm_euclid(const vector<double>& array)
{
    // host_math.cpp:67-72: sqrt of the dot product with itself
    return std::sqrt(device_dot_product(array, array));
}

double linalg::device_norm_inf(const vector<double>& array)
{
    double r = 0.0;
    check(lal_norm_inf_f64(ctx(), (int)array.get_size(), array.get_vec(), &r), "norm_inf");
    return r;
}

void linalg::device_jacobi_solve(const vector<double>& rhs,
                                 const vector<double>& diag,
                                 vector<double>&       x)
{
    check(lal_jacobi_solve_f64(ctx(), (int)rhs.get_size(), rhs.get_vec(), diag.get_vec(),
                               x.get_vec()),
          "jacobi_solve");
}

// ------------------------------------------------------------- primitives
// backend/device/device_primitives.h:38-47

template <typename T>
T linalg::device_find_min(const vector<T>&)
{
    unsupported(__func__);
    return T{};
}

template <typename T>
T linalg::device_find_max(const vector<T>&)
{
    unsupported(__func__);
    return T{};
}

template <typename T>
void linalg::device_exclusive_scan(vector<T>& x)
{
    if constexpr(std::is_same_v<T, int32_t>)
        check(lal_exclusive_scan_i32(ctx(), (int)x.get_size(), x.get_vec()), "exclusive_scan");
    else if constexpr(std::is_same_v<T, int64_t>)
        check(lal_exclusive_scan_i64(ctx(), (int)x.get_size(), x.get_vec()), "exclusive_scan");
    else
        unsupported(__func__);
}

template int32_t linalg::device_find_min<int32_t>(const vector<int32_t>& x);
template int64_t linalg::device_find_min<int64_t>(const vector<int64_t>& x);
template double  linalg::device_find_min<double>(const vector<double>& x);
template int32_t linalg::device_find_max<int32_t>(const vector<int32_t>& x);
template int64_t linalg::device_find_max<int64_t>(const vector<int64_t>& x);
template double  linalg::device_find_max<double>(const vector<double>& x);
template void    linalg::device_exclusive_scan<int32_t>(vector<int32_t>& x);
template void    linalg::device_exclusive_scan<int64_t>(vector<int64_t>& x);
template void    linalg::device_exclusive_scan<double>(vector<double>& x);

// --------------------------------------------------------- sparse products
// backend/device/device_matrix_vector.h:35-45, device_extract.h:34,
// device_csr2csc.h:33, device_csrgemm.h:37-44

void linalg::device_compute_residual(const csr_matrix<double>& A,
                                     const vector<double>&     x,
                                     const vector<double>&     b,
                                     vector<double>&           res)
{
    check(lal_residual_f64(ctx(), nullptr, A.get_m(), A.get_n(), A.get_nnz(), A.get_row_ptr(),
                           A.get_col_ind(), A.get_val(), x.get_vec(), b.get_vec(),
                           res.get_vec()),
          "compute_residual");
}

void linalg::free_csrmv_device_data(csrmv_descr* descr)
{
    lal_csrmv_descr d = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        auto                        it = g_csrmv.find(descr);
        if(it != g_csrmv.end())
        {
            d = it->second;
            g_csrmv.erase(it);
        }
    }
    if(d)
        check(lal_csrmv_descr_destroy(d), "free_csrmv_device_data");
}

void linalg::device_csrmv_analysis(const csr_matrix<double>& A,
                                   csrmv_algorithm,
                                   csrmv_descr* descr)
{
    free_csrmv_device_data(descr);
    lal_csrmv_descr d = nullptr;
    check(lal_csrmv_analysis(ctx(), A.get_m(), A.get_n(), A.get_nnz(), A.get_row_ptr(),
                             A.get_col_ind(), 0, &d),
          "csrmv_analysis");
    std::lock_guard<std::mutex> lock(g_mu);
    g_csrmv[descr] = d;
}

void linalg::device_csrmv_solve(double                    alpha,
                                const csr_matrix<double>& A,
                                const vector<double>&     x,
                                double                    beta,
                                vector<double>&           y,
                                csrmv_algorithm,
                                const csrmv_descr* descr)
{
    check(lal_csrmv_f64(ctx(), analysis_of(descr), A.get_m(), A.get_n(), A.get_nnz(), alpha,
                        A.get_row_ptr(), A.get_col_ind(), A.get_val(), x.get_vec(), beta,
                        y.get_vec()),
          "csrmv_solve");
}

void linalg::device_diagonal(const csr_matrix<double>& A, vector<double>& d)
{
    check(lal_diagonal_f64(ctx(), A.get_m(), A.get_row_ptr(), A.get_col_ind(), A.get_val(),
                           d.get_vec()),
          "diagonal");
}

void linalg::device_extract_lower_triangular_nnz(const csr_matrix<double>&,
                                                 csr_matrix<double>&,
                                                 int&)
{
    unsupported(__func__);
}
void linalg::device_extract_lower_triangular(const csr_matrix<double>&, csr_matrix<double>&)
{
    unsupported(__func__);
}
void linalg::device_extract_upper_triangular_nnz(const csr_matrix<double>&,
                                                 csr_matrix<double>&,
                                                 int&)
{
    unsupported(__func__);
}
void linalg::device_extract_upper_triangular(const csr_matrix<double>&, csr_matrix<double>&)
{
    unsupported(__func__);
}

void linalg::device_transpose_matrix(const csr_matrix<double>& A, csr_matrix<double>& transposeA)
{
    transposeA.resize(A.get_n(), A.get_m(), A.get_nnz());
    check(lal_transpose_f64(ctx(), A.get_m(), A.get_n(), A.get_nnz(), A.get_row_ptr(),
                            A.get_col_ind(), A.get_val(), transposeA.get_row_ptr(),
                            transposeA.get_col_ind(), transposeA.get_val()),
          "transpose");
}

void linalg::free_csrgemm_device_data(csrgemm_descr*) {}

void linalg::device_csrgemm_nnz(const csr_matrix<double>& A,
                                const csr_matrix<double>& B,
                                const csr_matrix<double>&,
                                csr_matrix<double>& C,
                                csrgemm_algorithm,
                                csrgemm_descr*)
{
    // as host_csrgemm_nnz (host_csrgemm.cpp:190-213): row pointer first, then
    // the column and value arrays sized by the product's nnz (D and beta are
    // not used by csr_matrix::multiply_by_matrix)
    C.resize(A.get_m(), B.get_n(), 0);
    int nnz_C = 0;
    check(lal_csrgemm_nnz(ctx(), A.get_m(), A.get_n(), B.get_n(), A.get_nnz(), A.get_row_ptr(),
                          A.get_col_ind(), B.get_nnz(), B.get_row_ptr(), B.get_col_ind(),
                          C.get_row_ptr(), &nnz_C),
          "csrgemm_nnz");
    C.resize(A.get_m(), B.get_n(), nnz_C);
}

void linalg::device_csrgemm_solve(double                    alpha,
                                  const csr_matrix<double>& A,
                                  const csr_matrix<double>& B,
                                  double,
                                  const csr_matrix<double>&,
                                  csr_matrix<double>& C,
                                  csrgemm_algorithm,
                                  const csrgemm_descr*)
{
    if(alpha != 1.0)
    {
        unsupported("device_csrgemm_solve with alpha != 1");
        return;
    }
    check(lal_csrgemm_solve_f64(ctx(), A.get_m(), A.get_n(), B.get_n(), A.get_nnz(),
                                A.get_row_ptr(), A.get_col_ind(), A.get_val(), B.get_nnz(),
                                B.get_row_ptr(), B.get_col_ind(), B.get_val(), C.get_row_ptr(),
                                C.get_col_ind(), C.get_val()),
          "csrgemm_solve");
}

// ------------------------------------------------------------ AMG setup
// backend/device/device_amg_strength.h:35, device_amg_aggregation.h:40-58

void linalg::device_compute_strong_connections(const csr_matrix<double>& A,
                                               double                    eps,
                                               vector<int>&              connections)
{
    check(lal_sa_strength_f64(ctx(), A.get_m(), A.get_nnz(), A.get_row_ptr(), A.get_col_ind(),
                              A.get_val(), eps, connections.get_vec()),
          "compute_strong_connections");
}

void linalg::device_compute_classical_strong_connections(const csr_matrix<double>&,
                                                         double,
                                                         csr_matrix<double>&,
                                                         vector<int>&)
{
    unsupported(__func__);
}

void linalg::device_initialize_pmis_state(const csr_matrix<double>& A,
                                          const vector<int>&        connections,
                                          vector<int>&              state,
                                          vector<int>&              hash)
{
    check(lal_pmis_init(ctx(), A.get_m(), A.get_nnz(), A.get_row_ptr(), connections.get_vec(),
                        state.get_vec(), hash.get_vec()),
          "initialize_pmis_state");
}

void linalg::device_find_maximum_distance_two_node(const csr_matrix<double>& A,
                                                   const vector<int>&        connections,
                                                   const vector<int>&        state,
                                                   const vector<int>&        hash,
                                                   vector<int64_t>&          aggregates,
                                                   vector<int>&              max_state,
                                                   bool&                     complete)
{
    int done = 1;
    check(lal_pmis_round(ctx(), A.get_m(), A.get_nnz(), A.get_row_ptr(), A.get_col_ind(),
                         connections.get_vec(), state.get_vec(), hash.get_vec(),
                         aggregates.get_vec(), max_state.get_vec(), &done),
          "find_maximum_distance_two_node");
    complete = done != 0;
}

void linalg::device_add_unassigned_nodes_to_closest_aggregation(
    const csr_matrix<double>& A,
    const vector<int>&        connections,
    const vector<int>&        state,
    vector<int64_t>&          aggregates,
    vector<int64_t>&,
    vector<int>& max_state)
{
    check(lal_pmis_join(ctx(), A.get_m(), A.get_nnz(), A.get_row_ptr(), A.get_col_ind(),
                        connections.get_vec(), state.get_vec(), aggregates.get_vec(),
                        max_state.get_vec()),
          "add_unassigned_nodes_to_closest_aggregation");
}

void linalg::device_compute_cfpoint_first_pass(const csr_matrix<double>&,
                                               const csr_matrix<double>&,
                                               vector<uint32_t>&)
{
    unsupported(__func__);
}

void linalg::device_compute_cfpoint_second_pass(const csr_matrix<double>&, vector<uint32_t>&)
{
    unsupported(__func__);
}

// ------------------------------------------------ outside the SA-AMG-PCG path

void linalg::free_csrgeam_device_data(csrgeam_descr*) {}
void linalg::device_csrgeam_nnz(const csr_matrix<double>&,
                                const csr_matrix<double>&,
                                csr_matrix<double>&,
                                csrgeam_algorithm,
                                csrgeam_descr*)
{
    unsupported(__func__);
}
void linalg::device_csrgeam_solve(double,
                                  const csr_matrix<double>&,
                                  double,
                                  const csr_matrix<double>&,
                                  csr_matrix<double>&,
                                  csrgeam_algorithm,
                                  const csrgeam_descr*)
{
    unsupported(__func__);
}

void linalg::device_csric0(csr_matrix<double>&, int*, int*)
{
    unsupported(__func__);
}
void linalg::free_csric0_device_data(csric0_descr*) {}
void linalg::device_csric0_analysis(const csr_matrix<double>&, csric0_descr*)
{
    unsupported(__func__);
}
void linalg::device_csric0_compute(csr_matrix<double>&, const csric0_descr*)
{
    unsupported(__func__);
}

void linalg::device_csrilu0(csr_matrix<double>&, int*, int*)
{
    unsupported(__func__);
}
void linalg::free_csrilu0_device_data(csrilu0_descr*) {}
void linalg::device_csrilu0_analysis(const csr_matrix<double>&, csrilu0_descr*)
{
    unsupported(__func__);
}
void linalg::device_csrilu0_compute(csr_matrix<double>&, const csrilu0_descr*)
{
    unsupported(__func__);
}

void linalg::free_csrtrsv_device_data(csrtrsv_descr*) {}
void linalg::device_csrtrsv_analysis(const csr_matrix<double>&,
                                     triangular_type,
                                     diagonal_type,
                                     csrtrsv_descr*)
{
    unsupported(__func__);
}
void linalg::device_csrtrsv_solve(const csr_matrix<double>&,
                                  const vector<double>&,
                                  vector<double>&,
                                  double,
                                  triangular_type,
                                  diagonal_type,
                                  const csrtrsv_descr*)
{
    unsupported(__func__);
}

void linalg::device_ruiz_scaling(vector<double>&, csr_matrix<double>&, vector<double>&, int, double)
{
    unsupported(__func__);
}
void linalg::device_symmetric_ruiz_scaling(vector<double>&, csr_matrix<double>&, int, double)
{
    unsupported(__func__);
}

void linalg::device_scale_diagonal(csr_matrix<double>&, double)
{
    unsupported(__func__);
}
void linalg::device_scale_by_inverse_diagonal(csr_matrix<double>&, const vector<double>&)
{
    unsupported(__func__);
}

void linalg::device_ssor_fill_lower_precond(const csr_matrix<double>&, csr_matrix<double>&, double)
{
    unsupported(__func__);
}
void linalg::device_ssor_fill_upper_precond(const csr_matrix<double>&, csr_matrix<double>&, double)
{
    unsupported(__func__);
}

template <typename T>
void linalg::device_partial_pivoting_algorithm(int,
                                               int,
                                               const vector<T>&,
                                               const vector<T>&,
                                               const vector<T>&,
                                               const vector<T>&,
                                               vector<T>&,
                                               pivoting_data<T>&)
{
    unsupported(__func__);
}

template <typename T>
void linalg::device_non_pivoting_algorithm(int,
                                           int,
                                           const vector<T>&,
                                           const vector<T>&,
                                           const vector<T>&,
                                           const vector<T>&,
                                           vector<T>&,
                                           non_pivoting_data<T>&)
{
    unsupported(__func__);
}

template void linalg::device_partial_pivoting_algorithm<float>(int,
                                                               int,
                                                               const vector<float>&,
                                                               const vector<float>&,
                                                               const vector<float>&,
                                                               const vector<float>&,
                                                               vector<float>&,
                                                               pivoting_data<float>&);
template void linalg::device_partial_pivoting_algorithm<double>(int,
                                                                int,
                                                                const vector<double>&,
                                                                const vector<double>&,
                                                                const vector<double>&,
                                                                const vector<double>&,
                                                                vector<double>&,
                                                                pivoting_data<double>&);
template void linalg::device_non_pivoting_algorithm<float>(int,
                                                           int,
                                                           const vector<float>&,
                                                           const vector<float>&,
                                                           const vector<float>&,
                                                           const vector<float>&,
                                                           vector<float>&,
                                                           non_pivoting_data<float>&);
template void linalg::device_non_pivoting_algorithm<double>(int,
                                                            int,
                                                            const vector<double>&,
                                                            const vector<double>&,
                                                            const vector<double>&,
                                                            const vector<double>&,
                                                            vector<double>&,
                                                            non_pivoting_data<double>&);
