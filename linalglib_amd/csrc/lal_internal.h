// Internal helpers shared by the gfx950 kernels and the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <vector>

#include "../../include/lal.h"

// Deterministic two-stage reductions use a fixed grid of this many blocks.
#define LAL_RED_BLOCKS 1024
#define LAL_RED_THREADS 256
// Device result slots in the context (dot/norm outputs, solver scalars).
#define LAL_NUM_SLOTS 64

struct lal_hierarchy_t;

// Devices the per-device static scratch table covers.
#define LAL_MAX_DEVICES 64

struct lal_context
{
    hipStream_t stream     = nullptr;
    bool        own_stream = false;
    int         device     = 0;
    int         num_cu     = 0;
    // reduction workspace: partials, a ticket counter per reduction slot
    double*   d_part   = nullptr; // LAL_RED_BLOCKS * 2
    unsigned* d_ticket = nullptr; // LAL_NUM_SLOTS
    double*   d_slots  = nullptr; // LAL_NUM_SLOTS device scalars
    double*   h_slots  = nullptr; // pinned mirror
    int*      d_flag   = nullptr;
    int*      h_flag   = nullptr;
    // hierarchies built on this context: lal_destroy detaches them, so that
    // one destroyed after its context no longer touches it
    std::vector<lal_hierarchy_t*> hierarchies;
};

namespace lal
{
    int  hip_status(hipError_t e, const char* what, const char* file, int line);
    void set_last_error(const char* msg);
}

#define LAL_HIP(call)                                                                  \
    do                                                                                 \
    {                                                                                  \
        hipError_t lal_e_ = (call);                                                    \
        if(lal_e_ != hipSuccess)                                                       \
            return lal::hip_status(lal_e_, #call, __FILE__, __LINE__);                 \
    } while(0)

#define LAL_LAUNCHED()                                                                 \
    do                                                                                 \
    {                                                                                  \
        hipError_t lal_e_ = hipGetLastError();                                         \
        if(lal_e_ != hipSuccess)                                                       \
            return lal::hip_status(lal_e_, "kernel launch", __FILE__, __LINE__);       \
    } while(0)

#define LAL_TRY(...)                                                                   \
    do                                                                                 \
    {                                                                                  \
        int lal_s_ = (__VA_ARGS__);                                                    \
        if(lal_s_ != LAL_OK)                                                           \
            return lal_s_;                                                             \
    } while(0)

#define LAL_REQUIRE(cond)                                                              \
    do                                                                                 \
    {                                                                                  \
        if(!(cond))                                                                    \
        {                                                                              \
            lal::set_last_error("invalid argument: " #cond);                           \
            return LAL_ERR_INVALID;                                                    \
        }                                                                              \
    } while(0)

namespace lal
{
    inline int grid_for(long long n, int threads, int cap = 1 << 30)
    {
        long long g = (n + threads - 1) / threads;
        if(g < 1)
            g = 1;
        if(g > cap)
            g = cap;
        return (int)g;
    }

    // Blocks are dispatched round-robin over the 8 XCDs of an MI355X (block b
    // runs on XCD b % 8).  This maps the blocks of one XCD onto one
    // contiguous range of logical tiles so that neighbouring rows (and the x
    // entries they share) stay in that XCD's L2.
    __device__ __forceinline__ int xcd_tile(int bid, int nb)
    {
        const int nx  = 8;
        int       per = nb / nx;
        int       rem = nb % nx;
        int       x   = bid % nx;
        int       k   = bid / nx;
        return x * per + (x < rem ? x : rem) + k;
    }

    // v + the value of v in another lane of the same quad, half row or row
    // (DPP control CTRL: quad_perm, row_half_mirror 0x141, row_mirror 0x140).
    template <int CTRL>
    __device__ __forceinline__ double dpp_add(double v)
    {
        const long long b  = __builtin_bit_cast(long long, v);
        const int       lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
        const int       hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
        return v + __builtin_bit_cast(double, (long long)(unsigned)lo | ((long long)hi << 32));
    }
    template <int CTRL>
    __device__ __forceinline__ float dpp_add(float v)
    {
        const int o = __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xF, 0xF, false);
        return v + __builtin_bit_cast(float, o);
    }

    // Sum over aligned groups of LANES (a power of two <= 64) lanes, left in
    // every lane of the group: lane ^ 1 and lane ^ 2 (quad permutations), the
    // half-row and row mirrors (each adds the other half's partial sum), then
    // exchanges with the lanes 16 and 32 away.
    template <int LANES, typename T>
    __device__ __forceinline__ T lane_sum(T v)
    {
        if(LANES >= 2)
            v = dpp_add<0xB1>(v); // quad_perm [1, 0, 3, 2]
        if(LANES >= 4)
            v = dpp_add<0x4E>(v); // quad_perm [2, 3, 0, 1]
        if(LANES >= 8)
            v = dpp_add<0x141>(v);
        if(LANES >= 16)
            v = dpp_add<0x140>(v);
        if(LANES >= 32)
            v += __shfl_xor(v, 16, 64);
        if(LANES >= 64)
            v += __shfl_xor(v, 32, 64);
        return v;
    }

    template <typename T>
    __device__ __forceinline__ T wave_sum(T v)
    {
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
            v += __shfl_xor(v, o, 64);
        return v;
    }

    template <typename T>
    __device__ __forceinline__ T wave_max(T v)
    {
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
        {
            T w = __shfl_xor(v, o, 64);
            v   = (w < v) ? v : w;
        }
        return v;
    }

    // Workgroup barrier that orders LDS accesses only: global loads issued
    // before it stay in flight (__syncthreads() waits for them).
    __device__ __forceinline__ void lds_barrier()
    {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
        __builtin_amdgcn_s_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
    }

    // Wavefront maximum that returns NaN when any lane holds one.
    template <typename T>
    __device__ __forceinline__ T wave_max_nan(T v)
    {
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
        {
            T w = __shfl_xor(v, o, 64);
            v   = (v != v || w < v) ? v : w;
        }
        return v;
    }

    // Internal launchers shared across translation units (device-resident
    // results; nothing is copied to the host).
    int dot_f64_dev(lal_context* c, int n, const double* x, const double* y, double* d_out,
                    int slot);
    int dot_f32_dev(lal_context* c, int n, const float* x, const float* y, double* d_out,
                    int slot);
    int norm_inf_f64_dev(lal_context* c, int n, const double* x, double* d_out, int slot);
    int scan_i32_dev(lal_context* c, int n, int32_t* x);
    int scan_i64_dev(lal_context* c, int n, int64_t* x);
    // Device address of the allocation-free scan's static scratch for
    // `device` (which must be current), looked up once per device.
    int noalloc_scratch_init(int device);
    int max_i64_dev(lal_context* c, int n, const int64_t* x, int64_t* h_out);
    // Clears the context pointer of every hierarchy built on c (lal_destroy).
    void detach_hierarchies(lal_context* c);
    // LAL_ERR_INVALID unless every ci[0..nnz) lies in [0, n) (blocks until known).
    int check_columns(lal_context* c, int nnz, int n, const int* ci);
}
