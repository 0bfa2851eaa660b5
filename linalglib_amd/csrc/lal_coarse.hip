// The coarse tail of the V-cycle in one workgroup.
//
// Below a few thousand nonzeros a level's kernels are pure launch latency, so
// the levels from `first` down to the coarsest run inside one 1024-lane
// workgroup, phase after phase with workgroup barriers: per level the first
// Jacobi sweep from zero (x = b/a_ii), r = b - A x, restriction; the coarsest
// solve; then per level on the way up x += P e and the post-sweep
// t = x + (b - A x)/a_ii.  This is the fast path with one pre- and one
// post-sweep (iterative_solvers/amg/amg.cpp:144-267 with n1 = n2 = 1); the
// arithmetic matches the per-level fast kernels up to summation order.
#include "lal_coarse.h"

namespace
{
    constexpr int CT = 1024; // lanes of the workgroup
    constexpr int CG = 16;   // lanes per row
    constexpr int RPP = CT / CG;

    template <typename T>
    __device__ __forceinline__ T group16(T v)
    {
#pragma unroll
        for(int o = CG / 2; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        return v;
    }

    // sum_j a_ij u_j for row `row` over the lanes of one 16-lane group
    template <typename T>
    __device__ __forceinline__ T row_sum(const int* __restrict__ rp, const int* __restrict__ ci,
                                         const T* __restrict__ v, const T* __restrict__ u,
                                         int row, bool active, int lane)
    {
        T acc = 0;
        if(active)
            for(int j = rp[row] + lane; j < rp[row + 1]; j += CG)
                acc += v[j] * u[ci[j]];
        return group16(acc);
    }

    template <typename T>
    __global__ void __launch_bounds__(CT)
        coarse_tail_kernel(const lal::coarse_level<T>* __restrict__ L, int nl, int kind,
                           const double* __restrict__ inv, T* __restrict__ result)
    {
        const int tid  = threadIdx.x;
        const int lane = tid & (CG - 1);
        const int slot = tid / CG;
        // down
        for(int k = 0; k + 1 < nl; k++)
        {
            const lal::coarse_level<T>& A = L[k];
            for(int i = tid; i < A.m; i += CT)
                A.x[i] = A.b[i] * A.dinv[i];
            __syncthreads();
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(A.rp, A.ci, A.v, A.x, row, act, lane);
                if(act && lane == 0)
                    A.r[row] = A.b[row] - s;
            }
            __syncthreads();
            const lal::coarse_level<T>& C = L[k + 1];
            for(int base = 0; base < C.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < C.m;
                const T    s   = row_sum(A.rrp, A.rci, A.rv, A.r, row, act, lane);
                if(act && lane == 0)
                    C.b[row] = s;
            }
            __syncthreads();
        }
        // coarsest
        {
            const lal::coarse_level<T>& C = L[nl - 1];
            if(kind == 0)
            {
                if(tid == 0)
                    C.x[0] = C.b[0] / C.v[0];
            }
            else
            {
                for(int base = 0; base < C.m; base += RPP)
                {
                    const int  row = base + slot;
                    const bool act = row < C.m;
                    double     acc = 0.0;
                    if(act)
                        for(int k = lane; k < C.m; k += CG)
                            acc = fma(inv[(size_t)row * C.m + k], (double)C.b[k], acc);
                    acc = group16(acc);
                    if(act && lane == 0)
                        C.x[row] = (T)acc;
                }
            }
            __syncthreads();
        }
        // up: x_k += P_k x_{k+1}; t_k = x_k + (b_k - A_k x_k)/a_ii
        const T* e = L[nl - 1].x;
        for(int k = nl - 2; k >= 0; k--)
        {
            const lal::coarse_level<T>& A = L[k];
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(A.prp, A.pci, A.pv, e, row, act, lane);
                if(act && lane == 0)
                    A.x[row] = s + A.x[row];
            }
            __syncthreads();
            T* out = (k == 0 && result) ? result : A.t;
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(A.rp, A.ci, A.v, A.x, row, act, lane);
                if(act && lane == 0)
                    out[row] = A.x[row] + (A.b[row] - s) * A.dinv[row];
            }
            __syncthreads();
            e = out;
        }
    }
}

namespace lal
{
    template <typename T>
    int coarse_tail(lal_context* c, const coarse_level<T>* d_levels, int nl, int kind,
                    const double* inv, T* result)
    {
        coarse_tail_kernel<T><<<1, CT, 0, c->stream>>>(d_levels, nl, kind, inv, result);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template int coarse_tail<double>(lal_context*, const coarse_level<double>*, int, int,
                                     const double*, double*);
    template int coarse_tail<float>(lal_context*, const coarse_level<float>*, int, int,
                                    const double*, float*);
}
