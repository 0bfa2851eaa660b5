// The coarse tail of the V-cycle in one workgroup, out of LDS.
//
// The last levels of a hierarchy hold a few thousand nonzeros: launched level
// by level their kernels are pure latency, and even inside one workgroup every
// phase would wait on a chain of dependent global loads.  So the setup packs
// those levels (rp, ci, values, 1/a_ii, P, R and, when it fits, the coarsest
// dense inverse) into one contiguous device image; the kernel copies the image
// into LDS with coalesced 16-byte loads, then runs every phase out of LDS with
// workgroup barriers between them: per level the first Jacobi sweep from zero
// (x = b/a_ii), r = b - A x and the restriction; the coarsest solve; then per
// level on the way up x += P e and the post-sweep t = x + (b - A x)/a_ii.
// This is the fast path with one pre- and one post-sweep
// (iterative_solvers/amg/amg.cpp:144-267 with n1 = n2 = 1); the arithmetic
// matches the per-level fast kernels up to summation order.
#include "lal_coarse.h"

namespace
{
    constexpr int CT  = 1024; // lanes of the workgroup
    constexpr int CG  = 16;   // lanes per row
    constexpr int RPP = CT / CG;

    template <typename T>
    __device__ __forceinline__ T group16(T v)
    {
#pragma unroll
        for(int o = CG / 2; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        return v;
    }

    // sum_j a_ij u_j of one row over the 16 lanes of a group (all in LDS)
    template <typename T>
    __device__ __forceinline__ T row_sum(const int* rp, const int* ci, const T* v, const T* u,
                                         int row, bool active, int lane)
    {
        T acc = 0;
        if(active)
            for(int j = rp[row] + lane; j < rp[row + 1]; j += CG)
                acc += v[j] * u[ci[j]];
        return group16(acc);
    }

    template <typename T>
    __device__ __forceinline__ T* at(int* lds, int off)
    {
        return reinterpret_cast<T*>(lds + off);
    }

    template <typename T>
    __global__ void __launch_bounds__(CT)
        coarse_tail_kernel(lal::tail_desc d, const int* __restrict__ image,
                           const T* __restrict__ b_in, const double* __restrict__ inv_g,
                           T* __restrict__ result)
    {
        extern __shared__ __attribute__((aligned(16))) int lds[];
        const int tid  = threadIdx.x;
        const int lane = tid & (CG - 1);
        const int slot = tid / CG;

        // image -> LDS (img_words is a multiple of 4, the image 16-byte aligned)
        {
            typedef int i4 __attribute__((ext_vector_type(4)));
            const i4* src = reinterpret_cast<const i4*>(image);
            i4*       dst = reinterpret_cast<i4*>(lds);
            for(int i = tid; i < d.img_words / 4; i += CT)
                dst[i] = src[i];
            T* b0 = at<T>(lds, d.L[0].b);
            for(int i = tid; i < d.L[0].m; i += CT)
                b0[i] = b_in[i];
        }
        __syncthreads();

        // down
        for(int k = 0; k + 1 < d.nl; k++)
        {
            const lal::tail_level& A  = d.L[k];
            const int*             rp = at<int>(lds, A.rp);
            const int*             ci = at<int>(lds, A.ci);
            const T*               v  = at<T>(lds, A.v);
            const T*               di = at<T>(lds, A.dinv);
            T*                     x  = at<T>(lds, A.x);
            T*                     b  = at<T>(lds, A.b);
            T*                     r  = at<T>(lds, A.r);
            for(int i = tid; i < A.m; i += CT)
                x[i] = b[i] * di[i];
            __syncthreads();
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(rp, ci, v, x, row, act, lane);
                if(act && lane == 0)
                    r[row] = b[row] - s;
            }
            __syncthreads();
            const lal::tail_level& C  = d.L[k + 1];
            T*                     bc = at<T>(lds, C.b);
            for(int base = 0; base < C.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < C.m;
                const T    s   = row_sum(at<int>(lds, A.rrp), at<int>(lds, A.rci),
                                         at<T>(lds, A.rv), r, row, act, lane);
                if(act && lane == 0)
                    bc[row] = s;
            }
            __syncthreads();
        }
        // coarsest
        {
            const lal::tail_level& C = d.L[d.nl - 1];
            T*                     x = at<T>(lds, C.x);
            const T*               b = at<T>(lds, C.b);
            if(d.kind == 0)
            {
                if(tid == 0)
                    x[0] = b[0] / at<T>(lds, C.v)[0];
            }
            else
            {
                const double* inv = d.inv >= 0 ? at<double>(lds, d.inv) : inv_g;
                for(int base = 0; base < C.m; base += RPP)
                {
                    const int  row = base + slot;
                    const bool act = row < C.m;
                    double     acc = 0.0;
                    if(act)
                        for(int k = lane; k < C.m; k += CG)
                            acc = fma(inv[(size_t)row * C.m + k], (double)b[k], acc);
                    acc = group16(acc);
                    if(act && lane == 0)
                        x[row] = (T)acc;
                }
            }
            __syncthreads();
        }
        // up: x_k += P_k e; t_k = x_k + (b_k - A_k x_k)/a_ii
        const T* e = at<T>(lds, d.L[d.nl - 1].x);
        for(int k = d.nl - 2; k >= 0; k--)
        {
            const lal::tail_level& A  = d.L[k];
            const int*             rp = at<int>(lds, A.rp);
            const int*             ci = at<int>(lds, A.ci);
            const T*               v  = at<T>(lds, A.v);
            const T*               di = at<T>(lds, A.dinv);
            T*                     x  = at<T>(lds, A.x);
            const T*               b  = at<T>(lds, A.b);
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(at<int>(lds, A.prp), at<int>(lds, A.pci),
                                         at<T>(lds, A.pv), e, row, act, lane);
                if(act && lane == 0)
                    x[row] = s + x[row];
            }
            __syncthreads();
            T* out = k == 0 ? result : at<T>(lds, A.t);
            for(int base = 0; base < A.m; base += RPP)
            {
                const int  row = base + slot;
                const bool act = row < A.m;
                const T    s   = row_sum(rp, ci, v, x, row, act, lane);
                if(act && lane == 0)
                    out[row] = x[row] + (b[row] - s) * di[row];
            }
            __syncthreads();
            e = out;
        }
    }
}

namespace lal
{
    template <typename T>
    int coarse_tail(lal_context* c, const tail_desc& d, const int* image, const T* b_in,
                    const double* inv, T* result)
    {
        if(d.nl < 2 || d.nl > TAIL_MAX_LEVELS || d.lds_words * 4 > TAIL_LDS_BYTES ||
           (d.img_words & 3) != 0 || (reinterpret_cast<uintptr_t>(image) & 15) != 0)
        {
            set_last_error("coarse tail: malformed level image");
            return LAL_ERR_INVALID;
        }
        coarse_tail_kernel<T><<<1, CT, (size_t)d.lds_words * 4, c->stream>>>(d, image, b_in, inv,
                                                                              result);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template int coarse_tail<double>(lal_context*, const tail_desc&, const int*, const double*,
                                     const double*, double*);
    template int coarse_tail<float>(lal_context*, const tail_desc&, const int*, const float*,
                                    const double*, float*);
}
