// BLAS-1 kernels: axpy family, deterministic reductions, exclusive scan.
//
// Arithmetic follows backend/host/host_axpy.cpp:34-97 and host_math.cpp:35-50
// (same special cases for alpha == 1, alpha == 0, beta == 0).  Reductions use
// a fixed grid and a fixed combination order, so results are bitwise
// repeatable run to run (no floating-point atomics).
#include "lal_internal.h"

#include <mutex>

// Scratch of lal_exclusive_scan_i64_noalloc: the scan's tile sums, one int64
// per 2048 entries (its tile), in a static device array so that the entry
// point allocates nothing.  Every GPU the library runs on holds its own copy
// until the process ends.  Host code never passes the array's name to a
// kernel (on the host it is a stand-in, not a device address): it uses the
// address hipGetSymbolAddress returned for that device (g_scratch).
static __device__ __align__(256) int64_t scratch[LAL_SCAN_NOALLOC_MAX_N / 2048];

static std::mutex g_scratch_mu;
static int64_t*   g_scratch[LAL_MAX_DEVICES] = {};

namespace
{
    // ---------------------------------------------------------------- axpy
    template <typename T>
    __global__ void __launch_bounds__(256)
        axpy_kernel(int n, T alpha, const T* __restrict__ x, T* __restrict__ y)
    {
        int i      = blockIdx.x * blockDim.x + threadIdx.x;
        int stride = gridDim.x * blockDim.x;
        if(alpha == (T)1)
            for(; i < n; i += stride)
                y[i] = x[i] + y[i];
        else
            for(; i < n; i += stride)
                y[i] = alpha * x[i] + y[i];
    }

    template <typename T>
    __global__ void __launch_bounds__(256)
        axpby_kernel(int n, T alpha, const T* __restrict__ x, T beta, T* __restrict__ y)
    {
        int i      = blockIdx.x * blockDim.x + threadIdx.x;
        int stride = gridDim.x * blockDim.x;
        if(alpha == (T)0)
            for(; i < n; i += stride)
                y[i] = beta * y[i];
        else if(beta == (T)0)
            for(; i < n; i += stride)
                y[i] = alpha * x[i];
        else
            for(; i < n; i += stride)
                y[i] = alpha * x[i] + beta * y[i];
    }

    template <typename T>
    __global__ void __launch_bounds__(256) axpbypgz_kernel(int n, T alpha,
                                                           const T* __restrict__ x, T beta,
                                                           const T* __restrict__ y, T gamma,
                                                           T* __restrict__ z)
    {
        int i      = blockIdx.x * blockDim.x + threadIdx.x;
        int stride = gridDim.x * blockDim.x;
        for(; i < n; i += stride)
            z[i] = alpha * x[i] + beta * y[i] + gamma * z[i];
    }

    // ----------------------------------------------------------- reductions
    // OP 0: sum x_i*y_i ; OP 1: max |x_i| ; OP 2: y_i = alpha x_i + y_i (the
    // axpy kernel's arithmetic) then sum y_i*z_i, which is lal_axpy then
    // lal_dot in one pass with the same partition and order.  Accumulation is
    // in double.
    template <typename T, int OP>
    __global__ void __launch_bounds__(LAL_RED_THREADS)
        reduce_kernel(int n, const T* __restrict__ x, const T* y, double* __restrict__ part,
                      unsigned* __restrict__ ticket, double* __restrict__ out, T alpha = 0,
                      T* yw = nullptr, const T* z = nullptr)
    {
        __shared__ double sh[LAL_RED_THREADS / 64];
        __shared__ bool   last;
        const int         tid = threadIdx.x;
        double            acc    = 0.0;
        const int         stride = gridDim.x * LAL_RED_THREADS;
        auto              add    = [&](T a, T b) {
            if(OP == 0)
                acc = acc + (double)a * (double)b;
            else
            {
                double t = fabs((double)a);
                acc      = (t < acc) ? acc : t;
            }
        };
        int i = blockIdx.x * LAL_RED_THREADS + tid;
        if(OP == 2)
        {
            // y = alpha x + y, then the products y_i * z_i; z may be y
            auto upd = [&](int k) {
                const T yn = alpha == (T)1 ? x[k] + yw[k] : alpha * x[k] + yw[k];
                yw[k]      = yn;
                return yn;
            };
            for(; i + 3 * stride < n; i += 4 * stride)
            {
                T a[4];
#pragma unroll
                for(int k = 0; k < 4; k++)
                    a[k] = upd(i + k * stride);
#pragma unroll
                for(int k = 0; k < 4; k++)
                    acc = acc + (double)a[k] * (double)(z == yw ? a[k] : z[i + k * stride]);
            }
            for(; i < n; i += stride)
            {
                const T a = upd(i);
                acc       = acc + (double)a * (double)(z == yw ? a : z[i]);
            }
        }
        else
        {
            for(; i + 3 * stride < n; i += 4 * stride)
            {
                T a[4], b[4];
#pragma unroll
                for(int k = 0; k < 4; k++)
                {
                    a[k] = x[i + k * stride];
                    b[k] = OP == 0 ? y[i + k * stride] : a[k];
                }
#pragma unroll
                for(int k = 0; k < 4; k++)
                    add(a[k], b[k]);
            }
            for(; i < n; i += stride)
                add(x[i], OP == 0 ? y[i] : x[i]);
        }
        constexpr bool SUM = OP != 1;
        acc = SUM ? lal::wave_sum(acc) : lal::wave_max(acc);
        if((tid & 63) == 0)
            sh[tid >> 6] = acc;
        __syncthreads();
        if(tid == 0)
        {
            double t = sh[0];
            for(int w = 1; w < LAL_RED_THREADS / 64; w++)
                t = SUM ? t + sh[w] : ((sh[w] < t) ? t : sh[w]);
            part[blockIdx.x] = t;
            __threadfence();
            unsigned k = atomicAdd(ticket, 1u);
            last       = (k == gridDim.x - 1);
        }
        __syncthreads();
        if(!last)
            return;
        __threadfence();
        double v = 0.0;
        for(int k = tid; k < (int)gridDim.x; k += LAL_RED_THREADS)
        {
            double p = __hip_atomic_load(&part[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            v        = SUM ? v + p : ((p < v) ? v : p);
        }
        v = SUM ? lal::wave_sum(v) : lal::wave_max(v);
        if((tid & 63) == 0)
            sh[tid >> 6] = v;
        __syncthreads();
        if(tid == 0)
        {
            double t = sh[0];
            for(int w = 1; w < LAL_RED_THREADS / 64; w++)
                t = SUM ? t + sh[w] : ((sh[w] < t) ? t : sh[w]);
            *out    = t;
            *ticket = 0u;
        }
    }

    // Dot products of at most SEQ_DOT entries are summed by one lane in index
    // order, s = s + x_i * y_i with both roundings, which is the host loop
    // (host_axpy.cpp:84-97) at one OpenMP thread bit for bit.  The wave first
    // stages both vectors in LDS so that the sequential chain reads LDS only.
    constexpr int SEQ_DOT = 1024;

    __global__ void __launch_bounds__(64)
        seq_dot_kernel(int n, const double* __restrict__ x, const double* __restrict__ y,
                       double* __restrict__ out)
    {
        __shared__ double sx[SEQ_DOT], sy[SEQ_DOT];
        for(int i = threadIdx.x; i < n; i += 64)
        {
            sx[i] = x[i];
            sy[i] = y[i];
        }
        __syncthreads();
        if(threadIdx.x != 0)
            return;
        double s = 0.0;
        for(int i = 0; i < n; i++)
            s = s + sx[i] * sy[i];
        *out = s;
    }

    template <typename T, int OP>
    int reduce_dev(lal_context* c, int n, const T* x, const T* y, double* d_out)
    {
        if(OP == 0 && sizeof(T) == 8 && n > 0 && n <= SEQ_DOT)
        {
            seq_dot_kernel<<<1, 64, 0, c->stream>>>(n, (const double*)x, (const double*)y,
                                                    d_out);
            LAL_LAUNCHED();
            return LAL_OK;
        }
        int g = n <= 0 ? 1 : lal::grid_for(n, LAL_RED_THREADS, LAL_RED_BLOCKS);
        reduce_kernel<T, OP>
            <<<g, LAL_RED_THREADS, 0, c->stream>>>(n, x, y, c->d_part, c->d_ticket, d_out);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    // ------------------------------------------------------------- scan
    constexpr int SCAN_THREADS = 256;
    constexpr int SCAN_ITEMS   = 8;
    constexpr int SCAN_TILE    = SCAN_THREADS * SCAN_ITEMS;

    // Exclusive scan of one value per thread across the block; returns the
    // thread's exclusive prefix and writes the block total.
    template <typename T>
    __device__ T block_exclusive_scan(T v, T* total)
    {
        __shared__ T wsum[SCAN_THREADS / 64];
        const int    lane = threadIdx.x & 63;
        const int    wid  = threadIdx.x >> 6;
        T            inc  = v;
#pragma unroll
        for(int o = 1; o < 64; o <<= 1)
        {
            T t = __shfl_up(inc, o, 64);
            if(lane >= o)
                inc += t;
        }
        if(lane == 63)
            wsum[wid] = inc;
        __syncthreads();
        T off = 0, tot = 0;
        for(int w = 0; w < SCAN_THREADS / 64; w++)
        {
            if(w < wid)
                off += wsum[w];
            tot += wsum[w];
        }
        __syncthreads();
        *total = tot;
        return off + inc - v;
    }

    template <typename T>
    __global__ void __launch_bounds__(SCAN_THREADS)
        scan_tile_sums(int n, const T* __restrict__ x, T* __restrict__ sums)
    {
        int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
        T   s    = 0;
#pragma unroll
        for(int k = 0; k < SCAN_ITEMS; k++)
            if(base + k < n)
                s += x[base + k];
        T tot;
        (void)block_exclusive_scan(s, &tot);
        if(threadIdx.x == 0)
            sums[blockIdx.x] = tot;
    }

    // Single block: exclusive scan of the tile sums in chunks.
    template <typename T>
    __global__ void __launch_bounds__(SCAN_THREADS) scan_sums(int ntiles, T* __restrict__ sums)
    {
        T carry = 0;
        for(int base = 0; base < ntiles; base += SCAN_THREADS)
        {
            int i = base + threadIdx.x;
            T   v = i < ntiles ? sums[i] : 0;
            T   tot;
            T   e = block_exclusive_scan(v, &tot);
            if(i < ntiles)
                sums[i] = carry + e;
            carry += tot;
            __syncthreads();
        }
    }

    template <typename T>
    __global__ void __launch_bounds__(SCAN_THREADS)
        scan_apply(int n, T* __restrict__ x, const T* __restrict__ sums)
    {
        int base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
        T   v[SCAN_ITEMS];
        T   s = 0;
#pragma unroll
        for(int k = 0; k < SCAN_ITEMS; k++)
        {
            v[k] = base + k < n ? x[base + k] : 0;
            s += v[k];
        }
        T tot;
        T e = block_exclusive_scan(s, &tot) + sums[blockIdx.x];
#pragma unroll
        for(int k = 0; k < SCAN_ITEMS; k++)
        {
            if(base + k < n)
                x[base + k] = e;
            e += v[k];
        }
    }

    static_assert(SCAN_TILE == 2048, "the no-allocation scratch holds one sum per 2048 entries");

    // The three passes of the scan with the tile sums in `sums` (ntiles
    // entries, every one written by the first pass before it is read).
    template <typename T>
    int scan_passes(lal_context* c, int n, T* x, T* sums)
    {
        const int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
        scan_tile_sums<T><<<ntiles, SCAN_THREADS, 0, c->stream>>>(n, x, sums);
        LAL_LAUNCHED();
        scan_sums<T><<<1, SCAN_THREADS, 0, c->stream>>>(ntiles, sums);
        LAL_LAUNCHED();
        scan_apply<T><<<ntiles, SCAN_THREADS, 0, c->stream>>>(n, x, sums);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int scan_dev(lal_context* c, int n, T* x)
    {
        if(n <= 0)
            return LAL_OK;
        int ntiles = (n + SCAN_TILE - 1) / SCAN_TILE;
        T*  sums   = nullptr;
        LAL_HIP(hipMallocAsync((void**)&sums, sizeof(T) * ntiles, c->stream));
        const int st = scan_passes<T>(c, n, x, sums);
        LAL_HIP(hipFreeAsync(sums, c->stream));
        return st;
    }

    __global__ void __launch_bounds__(256)
        max_i64_kernel(int n, const int64_t* __restrict__ x, unsigned long long* out)
    {
        long long m = LLONG_MIN;
        for(int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
            m = x[i] > m ? x[i] : m;
#pragma unroll
        for(int o = 32; o > 0; o >>= 1)
        {
            long long t = __shfl_xor(m, o, 64);
            m           = t > m ? t : m;
        }
        if((threadIdx.x & 63) == 0)
            atomicMax((long long*)out, m);
    }

    template <typename T>
    int axpy_t(lal_context* c, int n, T alpha, const T* x, T* y)
    {
        LAL_REQUIRE(c != nullptr && n >= 0);
        if(n == 0)
            return LAL_OK;
        LAL_REQUIRE(x != nullptr && y != nullptr);
        axpy_kernel<T><<<lal::grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, alpha, x, y);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename T>
    int axpby_t(lal_context* c, int n, T alpha, const T* x, T beta, T* y)
    {
        LAL_REQUIRE(c != nullptr && n >= 0);
        if(n == 0)
            return LAL_OK;
        LAL_REQUIRE(y != nullptr && (x != nullptr || alpha == (T)0));
        axpby_kernel<T>
            <<<lal::grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, alpha, x, beta, y);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

namespace lal
{
    int dot_f64_dev(lal_context* c, int n, const double* x, const double* y, double* d_out, int)
    {
        return reduce_dev<double, 0>(c, n, x, y, d_out);
    }
    int dot_f32_dev(lal_context* c, int n, const float* x, const float* y, double* d_out, int)
    {
        return reduce_dev<float, 0>(c, n, x, y, d_out);
    }
    int norm_inf_f64_dev(lal_context* c, int n, const double* x, double* d_out, int)
    {
        return reduce_dev<double, 1>(c, n, x, x, d_out);
    }
    int scan_i32_dev(lal_context* c, int n, int32_t* x)
    {
        return scan_dev<int32_t>(c, n, x);
    }
    int scan_i64_dev(lal_context* c, int n, int64_t* x)
    {
        return scan_dev<int64_t>(c, n, x);
    }

    // Called by lal_create with `device` current: the device address of that
    // device's scratch array, looked up once per device.
    int noalloc_scratch_init(int device)
    {
        if(device < 0 || device >= LAL_MAX_DEVICES)
        {
            set_last_error("device ordinal beyond LAL_MAX_DEVICES");
            return LAL_ERR_NO_DEVICE;
        }
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if(g_scratch[device] != nullptr)
            return LAL_OK;
        void* p = nullptr;
        LAL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if(p == nullptr)
        {
            set_last_error("hipGetSymbolAddress returned no address for the scan scratch");
            return LAL_ERR_HIP;
        }
        g_scratch[device] = static_cast<int64_t*>(p);
        return LAL_OK;
    }
    int max_i64_dev(lal_context* c, int n, const int64_t* x, int64_t* h_out)
    {
        long long init = LLONG_MIN;
        auto*     slot = reinterpret_cast<unsigned long long*>(c->d_slots + (LAL_NUM_SLOTS - 1));
        LAL_HIP(hipMemcpyAsync(slot, &init, sizeof(init), hipMemcpyHostToDevice, c->stream));
        if(n > 0)
        {
            max_i64_kernel<<<lal::grid_for(n, 256, 2048), 256, 0, c->stream>>>(n, x, slot);
            LAL_LAUNCHED();
        }
        LAL_HIP(hipMemcpyAsync(h_out, slot, sizeof(long long), hipMemcpyDeviceToHost, c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        return LAL_OK;
    }
}

extern "C" int lal_axpy_f64(lal_handle c, int n, double alpha, const double* x, double* y)
{
    return axpy_t(c, n, alpha, x, y);
}
extern "C" int lal_axpy_f32(lal_handle c, int n, float alpha, const float* x, float* y)
{
    return axpy_t(c, n, alpha, x, y);
}
extern "C" int lal_axpby_f64(lal_handle c, int n, double alpha, const double* x, double beta,
                             double* y)
{
    return axpby_t(c, n, alpha, x, beta, y);
}
extern "C" int lal_axpby_f32(lal_handle c, int n, float alpha, const float* x, float beta,
                             float* y)
{
    return axpby_t(c, n, alpha, x, beta, y);
}
extern "C" int lal_axpbypgz_f64(lal_handle c, int n, double alpha, const double* x, double beta,
                                const double* y, double gamma, double* z)
{
    LAL_REQUIRE(c != nullptr && n >= 0);
    if(n == 0)
        return LAL_OK;
    LAL_REQUIRE(x != nullptr && y != nullptr && z != nullptr);
    axpbypgz_kernel<double>
        <<<lal::grid_for(n, 256, 16384), 256, 0, c->stream>>>(n, alpha, x, beta, y, gamma, z);
    LAL_LAUNCHED();
    return LAL_OK;
}

namespace
{
    template <typename T, int OP>
    int reduce_host(lal_context* c, int n, const T* x, const T* y, double* result)
    {
        LAL_REQUIRE(c != nullptr && n >= 0 && result != nullptr);
        LAL_REQUIRE(n == 0 || (x != nullptr && y != nullptr));
        LAL_TRY((reduce_dev<T, OP>(c, n, x, y, c->d_slots)));
        LAL_HIP(hipMemcpyAsync(c->h_slots, c->d_slots, sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
        LAL_HIP(hipStreamSynchronize(c->stream));
        *result = c->h_slots[0];
        return LAL_OK;
    }
}

extern "C" int lal_dot_f64(lal_handle c, int n, const double* x, const double* y, double* r)
{
    return reduce_host<double, 0>(c, n, x, y, r);
}

extern "C" int lal_axpy_dot_f64(lal_handle c, int n, double alpha, const double* x, double* y,
                                const double* z, double* result)
{
    LAL_REQUIRE(c != nullptr && n >= 0 && result != nullptr);
    LAL_REQUIRE(n == 0 || (x != nullptr && y != nullptr && z != nullptr));
    if(n > 0 && n <= SEQ_DOT)
    {
        // short vectors: the axpy, then the dot of lal_dot_f64's short path
        LAL_TRY(axpy_t<double>(c, n, alpha, x, y));
        return reduce_host<double, 0>(c, n, y, z, result);
    }
    const int g = n <= 0 ? 1 : lal::grid_for(n, LAL_RED_THREADS, LAL_RED_BLOCKS);
    reduce_kernel<double, 2><<<g, LAL_RED_THREADS, 0, c->stream>>>(
        n, x, y, c->d_part, c->d_ticket, c->d_slots, alpha, y, z);
    LAL_LAUNCHED();
    LAL_HIP(hipMemcpyAsync(c->h_slots, c->d_slots, sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    LAL_HIP(hipStreamSynchronize(c->stream));
    *result = c->h_slots[0];
    return LAL_OK;
}
extern "C" int lal_dot_f32(lal_handle c, int n, const float* x, const float* y, double* r)
{
    return reduce_host<float, 0>(c, n, x, y, r);
}
extern "C" int lal_norm_inf_f64(lal_handle c, int n, const double* x, double* r)
{
    return reduce_host<double, 1>(c, n, x, x, r);
}
extern "C" int lal_norm_inf_f32(lal_handle c, int n, const float* x, double* r)
{
    return reduce_host<float, 1>(c, n, x, x, r);
}
extern "C" int lal_exclusive_scan_i32(lal_handle c, int n, int32_t* x)
{
    LAL_REQUIRE(c != nullptr && n >= 0 && (n == 0 || x != nullptr));
    return lal::scan_i32_dev(c, n, x);
}
extern "C" int lal_exclusive_scan_i64(lal_handle c, int n, int64_t* x)
{
    LAL_REQUIRE(c != nullptr && n >= 0 && (n == 0 || x != nullptr));
    return lal::scan_i64_dev(c, n, x);
}

extern "C" int lal_exclusive_scan_i64_noalloc(lal_handle c, int n, int64_t* x)
{
    LAL_REQUIRE(c != nullptr && n >= 0 && (n == 0 || x != nullptr));
    if(n > LAL_SCAN_NOALLOC_MAX_N)
    {
        char buf[160];
        snprintf(buf, sizeof(buf),
                 "lal_exclusive_scan_i64_noalloc: n = %d exceeds the static scratch limit "
                 "LAL_SCAN_NOALLOC_MAX_N = %d (use lal_exclusive_scan_i64)",
                 n, LAL_SCAN_NOALLOC_MAX_N);
        lal::set_last_error(buf);
        return LAL_ERR_INVALID;
    }
    if(n == 0)
        return LAL_OK;
    int64_t* sums = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if(c->device >= 0 && c->device < LAL_MAX_DEVICES)
            sums = g_scratch[c->device];
    }
    if(sums == nullptr)
    {
        lal::set_last_error("lal_exclusive_scan_i64_noalloc: no scratch address for the "
                            "handle's device (lal_create did not set it up)");
        return LAL_ERR_INVALID;
    }
    // the first pass writes every tile sum the next two read: nothing to clear
    return scan_passes<int64_t>(c, n, x, sums);
}

