// Context, memory and elementwise utilities of the C ABI.
#include "lal_internal.h"

#include <cstring>
#include <mutex>
#include <string>

namespace lal
{
    static thread_local std::string g_last_error;

    void set_last_error(const char* msg)
    {
        g_last_error = msg;
    }

    int hip_status(hipError_t e, const char* what, const char* file, int line)
    {
        char buf[512];
        snprintf(buf, sizeof(buf), "%s failed: %s (%s:%d)", what, hipGetErrorString(e), file,
                 line);
        g_last_error = buf;
        if(e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
            return LAL_ERR_OOM;
        if(e == hipErrorNoDevice || e == hipErrorInvalidDevice)
            return LAL_ERR_NO_DEVICE;
        return LAL_ERR_HIP;
    }
}

extern "C" const char* lal_last_error(void)
{
    return lal::g_last_error.c_str();
}

extern "C" const char* lal_status_string(int s)
{
    switch(s)
    {
    case LAL_OK: return "ok";
    case LAL_ERR_INVALID: return "invalid argument";
    case LAL_ERR_OOM: return "out of device memory";
    case LAL_ERR_HIP: return "HIP runtime error";
    case LAL_ERR_NOT_CONVERGED: return "not converged";
    case LAL_ERR_NAN: return "NaN in residual";
    case LAL_ERR_NO_DEVICE: return "no usable device";
    case LAL_ERR_UNSUPPORTED: return "unsupported input";
    default: return "unknown status";
    }
}

extern "C" int lal_create(lal_handle* handle)
{
    LAL_REQUIRE(handle != nullptr);
    *handle = nullptr;
    int ndev = 0;
    if(hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1)
    {
        lal::set_last_error("no HIP device");
        return LAL_ERR_NO_DEVICE;
    }
    lal_context* c = new lal_context;
    LAL_HIP(hipGetDevice(&c->device));
    hipDeviceProp_t prop;
    LAL_HIP(hipGetDeviceProperties(&prop, c->device));
    if(std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
    {
        char buf[256];
        snprintf(buf, sizeof(buf), "device is %s, this library is built for gfx950",
                 prop.gcnArchName);
        lal::set_last_error(buf);
        delete c;
        return LAL_ERR_NO_DEVICE;
    }
    c->num_cu = prop.multiProcessorCount;
    {
        const int st = lal::noalloc_scratch_init(c->device);
        if(st != LAL_OK)
        {
            delete c;
            return st;
        }
    }
    LAL_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    c->own_stream = true;
    LAL_HIP(hipMalloc(&c->d_part, sizeof(double) * LAL_RED_BLOCKS * 2));
    LAL_HIP(hipMalloc(&c->d_ticket, sizeof(unsigned) * LAL_NUM_SLOTS));
    LAL_HIP(hipMalloc(&c->d_slots, sizeof(double) * LAL_NUM_SLOTS));
    LAL_HIP(hipMalloc(&c->d_flag, sizeof(int) * 16));
    LAL_HIP(hipHostMalloc(&c->h_slots, sizeof(double) * LAL_NUM_SLOTS, hipHostMallocDefault));
    LAL_HIP(hipHostMalloc(&c->h_flag, sizeof(int) * 16, hipHostMallocDefault));
    LAL_HIP(hipMemsetAsync(c->d_ticket, 0, sizeof(unsigned) * LAL_NUM_SLOTS, c->stream));
    LAL_HIP(hipMemsetAsync(c->d_slots, 0, sizeof(double) * LAL_NUM_SLOTS, c->stream));
    LAL_HIP(hipMemsetAsync(c->d_flag, 0, sizeof(int) * 16, c->stream));
    LAL_HIP(hipStreamSynchronize(c->stream));
    *handle = c;
    return LAL_OK;
}

extern "C" int lal_destroy(lal_handle c)
{
    if(!c)
        return LAL_OK;
    if(c->stream)
        (void)hipStreamSynchronize(c->stream);
    lal::detach_hierarchies(c);
    (void)hipFree(c->d_part);
    (void)hipFree(c->d_ticket);
    (void)hipFree(c->d_slots);
    (void)hipFree(c->d_flag);
    (void)hipHostFree(c->h_slots);
    (void)hipHostFree(c->h_flag);
    if(c->own_stream)
        (void)hipStreamDestroy(c->stream);
    delete c;
    return LAL_OK;
}

extern "C" int lal_set_stream(lal_handle c, void* stream)
{
    LAL_REQUIRE(c != nullptr);
    LAL_HIP(hipStreamSynchronize(c->stream));
    if(c->own_stream)
        LAL_HIP(hipStreamDestroy(c->stream));
    c->stream     = static_cast<hipStream_t>(stream);
    c->own_stream = false;
    return LAL_OK;
}

extern "C" int lal_get_stream(lal_handle c, void** stream)
{
    LAL_REQUIRE(c != nullptr && stream != nullptr);
    *stream = c->stream;
    return LAL_OK;
}

extern "C" int lal_sync(lal_handle c)
{
    LAL_REQUIRE(c != nullptr);
    LAL_HIP(hipStreamSynchronize(c->stream));
    return LAL_OK;
}

extern "C" int lal_alloc(void** ptr, size_t bytes)
{
    LAL_REQUIRE(ptr != nullptr);
    *ptr = nullptr;
    if(bytes == 0)
        return LAL_OK;
    LAL_HIP(hipMalloc(ptr, bytes));
    return LAL_OK;
}

extern "C" int lal_free(void* ptr)
{
    if(ptr)
        LAL_HIP(hipFree(ptr));
    return LAL_OK;
}

extern "C" int lal_memcpy_h2d(lal_handle c, void* dst, const void* src, size_t bytes)
{
    LAL_REQUIRE(c != nullptr);
    if(bytes == 0)
        return LAL_OK;
    LAL_REQUIRE(dst != nullptr && src != nullptr);
    LAL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    LAL_HIP(hipStreamSynchronize(c->stream));
    return LAL_OK;
}

extern "C" int lal_memcpy_d2h(lal_handle c, void* dst, const void* src, size_t bytes)
{
    LAL_REQUIRE(c != nullptr);
    if(bytes == 0)
        return LAL_OK;
    LAL_REQUIRE(dst != nullptr && src != nullptr);
    LAL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    LAL_HIP(hipStreamSynchronize(c->stream));
    return LAL_OK;
}

extern "C" int lal_memcpy_d2d(lal_handle c, void* dst, const void* src, size_t bytes)
{
    LAL_REQUIRE(c != nullptr);
    if(bytes == 0)
        return LAL_OK;
    LAL_REQUIRE(dst != nullptr && src != nullptr);
    LAL_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return LAL_OK;
}

namespace
{
    template <typename T>
    __global__ void __launch_bounds__(256) fill_kernel(T* __restrict__ x, size_t n, T v)
    {
        size_t i      = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        size_t stride = (size_t)gridDim.x * blockDim.x;
        for(; i < n; i += stride)
            x[i] = v;
    }

    template <typename S, typename D>
    __global__ void __launch_bounds__(256)
        convert_kernel(const S* __restrict__ x, D* __restrict__ y, size_t n)
    {
        size_t i      = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        size_t stride = (size_t)gridDim.x * blockDim.x;
        for(; i < n; i += stride)
            y[i] = static_cast<D>(x[i]);
    }

    template <typename T>
    int fill(lal_context* c, T* x, size_t n, T v)
    {
        LAL_REQUIRE(c != nullptr);
        if(n == 0)
            return LAL_OK;
        LAL_REQUIRE(x != nullptr);
        int g = lal::grid_for((long long)n, 256, 8192);
        fill_kernel<T><<<g, 256, 0, c->stream>>>(x, n, v);
        LAL_LAUNCHED();
        return LAL_OK;
    }

    template <typename S, typename D>
    int convert(lal_context* c, const S* x, D* y, size_t n)
    {
        LAL_REQUIRE(c != nullptr);
        if(n == 0)
            return LAL_OK;
        LAL_REQUIRE(x != nullptr && y != nullptr);
        int g = lal::grid_for((long long)n, 256, 8192);
        convert_kernel<S, D><<<g, 256, 0, c->stream>>>(x, y, n);
        LAL_LAUNCHED();
        return LAL_OK;
    }
}

extern "C" int lal_fill_f64(lal_handle c, double* x, size_t n, double v)
{
    return fill(c, x, n, v);
}
extern "C" int lal_fill_f32(lal_handle c, float* x, size_t n, float v)
{
    return fill(c, x, n, v);
}
extern "C" int lal_fill_i32(lal_handle c, int32_t* x, size_t n, int32_t v)
{
    return fill(c, x, n, v);
}
extern "C" int lal_fill_i64(lal_handle c, int64_t* x, size_t n, int64_t v)
{
    return fill(c, x, n, v);
}
extern "C" int lal_convert_f64_f32(lal_handle c, const double* x, float* y, size_t n)
{
    return convert(c, x, y, n);
}
extern "C" int lal_convert_f32_f64(lal_handle c, const float* x, double* y, size_t n)
{
    return convert(c, x, y, n);
}
