// Runtime part of the C ABI: errors, device query, memory, pin kernel,
// copy-peak probe.
#include "clh_internal.h"

#include <cstring>
#include <string>

namespace clh {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg)
{
    g_last_error = msg;
    return status;
}

int hip_status(hipError_t e, const char* what)
{
    if (e == hipSuccess) return CLH_OK;
    return fail(CLH_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

}  // namespace clh

using namespace clh;

extern "C" const char* clh_last_error(void) { return g_last_error.c_str(); }

extern "C" int clh_device_count(int* count)
{
    if (!count) return fail(CLH_ERR_INVALID, "clh_device_count: null pointer");
    *count = 0;
    const hipError_t e = hipGetDeviceCount(count);
    if (e == hipErrorNoDevice) {
        *count = 0;
        (void)hipGetLastError();
        return CLH_OK;
    }
    return hip_status(e, "hipGetDeviceCount");
}

extern "C" int clh_get_device_info(int device, clh_device_info* out)
{
    if (!out) return fail(CLH_ERR_INVALID, "clh_get_device_info: null pointer");
    hipDeviceProp_t p;
    CLH_HIP(hipGetDeviceProperties(&p, device));
    std::memset(out, 0, sizeof(*out));
    std::strncpy(out->name, p.name, sizeof(out->name) - 1);
    std::strncpy(out->arch, p.gcnArchName, sizeof(out->arch) - 1);
    out->compute_units  = p.multiProcessorCount;
    out->clock_khz      = p.clockRate;
    out->mem_clock_khz  = p.memoryClockRate;
    out->mem_bus_width  = p.memoryBusWidth;
    out->total_mem      = p.totalGlobalMem;
    out->l2_cache_bytes = (size_t)p.l2CacheSize;
    return CLH_OK;
}

extern "C" int clh_alloc(void** ptr, size_t bytes)
{
    if (!ptr) return fail(CLH_ERR_INVALID, "clh_alloc: null pointer");
    *ptr = nullptr;
    if (bytes == 0) return CLH_OK;
    const hipError_t e = hipMalloc(ptr, bytes);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(CLH_ERR_ALLOC, std::string("hipMalloc(") + std::to_string(bytes) +
                                       " bytes): " + hipGetErrorString(e));
    }
    return CLH_OK;
}

extern "C" int clh_free(void* ptr)
{
    if (!ptr) return CLH_OK;
    return hip_status(hipFree(ptr), "hipFree");
}

extern "C" int clh_memcpy(void* dst, const void* src, size_t bytes, int kind)
{
    if (bytes == 0) return CLH_OK;
    if (!dst || !src) return fail(CLH_ERR_INVALID, "clh_memcpy: null pointer");
    hipMemcpyKind k;
    switch (kind) {
        case CLH_H2H: k = hipMemcpyHostToHost; break;
        case CLH_H2D: k = hipMemcpyHostToDevice; break;
        case CLH_D2H: k = hipMemcpyDeviceToHost; break;
        case CLH_D2D: k = hipMemcpyDeviceToDevice; break;
        default: return fail(CLH_ERR_INVALID, "clh_memcpy: unknown copy kind");
    }
    CLH_HIP(hipMemcpy(dst, src, bytes, k));
    if (kind == CLH_D2D) CLH_HIP(hipStreamSynchronize(0));
    return CLH_OK;
}

extern "C" int clh_memset(void* ptr, int value, size_t bytes)
{
    if (bytes == 0) return CLH_OK;
    if (!ptr) return fail(CLH_ERR_INVALID, "clh_memset: null pointer");
    CLH_HIP(hipMemset(ptr, value, bytes));
    CLH_HIP(hipStreamSynchronize(0));
    return CLH_OK;
}

extern "C" int clh_synchronize(void) { return hip_status(hipDeviceSynchronize(), "hipDeviceSynchronize"); }

// pos[idx[i]] = target[i], vel[idx[i]] = 0.
__global__ void __launch_bounds__(256)
k_apply_pins(float* __restrict__ pos, float* __restrict__ vel, const int* __restrict__ idx,
             const float* __restrict__ target, int count)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const size_t v = (size_t)idx[i] * 3;
    pos[v + 0] = target[(size_t)i * 3 + 0];
    pos[v + 1] = target[(size_t)i * 3 + 1];
    pos[v + 2] = target[(size_t)i * 3 + 2];
    if (vel) {
        vel[v + 0] = 0.0f;
        vel[v + 1] = 0.0f;
        vel[v + 2] = 0.0f;
    }
}

extern "C" int clh_apply_constraints(float* d_pos, float* d_vel, const int* d_indices,
                                     const float* d_target, int num_pinned)
{
    if (num_pinned == 0) return CLH_OK;
    if (num_pinned < 0 || !d_pos || !d_indices || !d_target)
        return fail(CLH_ERR_INVALID, "clh_apply_constraints: bad arguments");
    hipLaunchKernelGGL(k_apply_pins, dim3((num_pinned + 255) / 256), dim3(256), 0, 0, d_pos,
                       d_vel, d_indices, d_target, num_pinned);
    return hip_status(hipGetLastError(), "k_apply_pins launch");
}

// One 16-byte element per thread; the tail (bytes % 16) is copied bytewise.
__global__ void __launch_bounds__(256)
k_copy16(uint4* __restrict__ dst, const uint4* __restrict__ src, size_t n16,
         unsigned char* __restrict__ dtail, const unsigned char* __restrict__ stail, int ntail)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n16) dst[i] = src[i];
    if (blockIdx.x == 0 && (int)threadIdx.x < ntail) dtail[threadIdx.x] = stail[threadIdx.x];
}

extern "C" int clh_measure_copy(void* d_dst, const void* d_src, size_t bytes, int num_launches,
                                float* avg_ms)
{
    if (!d_dst || !d_src || !avg_ms || num_launches <= 0 || bytes < 16)
        return fail(CLH_ERR_INVALID, "clh_measure_copy: bad arguments");
    if (((uintptr_t)d_dst | (uintptr_t)d_src) & 15)
        return fail(CLH_ERR_INVALID, "clh_measure_copy: buffers must be 16-byte aligned");
    const size_t n16 = bytes / 16;
    const int ntail = (int)(bytes % 16);
    const unsigned blocks = (unsigned)((n16 + 255) / 256);
    auto launch = [&]() {
        hipLaunchKernelGGL(k_copy16, dim3(blocks), dim3(256), 0, 0, (uint4*)d_dst,
                           (const uint4*)d_src, n16, (unsigned char*)d_dst + n16 * 16,
                           (const unsigned char*)d_src + n16 * 16, ntail);
    };
    launch();  // warm-up
    CLH_HIP(hipGetLastError());
    hipEvent_t a, b;
    CLH_HIP(hipEventCreate(&a));
    CLH_HIP(hipEventCreate(&b));
    CLH_HIP(hipEventRecord(a, 0));
    for (int i = 0; i < num_launches; ++i) launch();
    CLH_HIP(hipEventRecord(b, 0));
    CLH_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    CLH_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = ms / (float)num_launches;
    return CLH_OK;
}
