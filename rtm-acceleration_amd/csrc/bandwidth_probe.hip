// bandwidth_probe.hip -- device copy bandwidth (the reference's GPU probe,
// src/tools/bandwidth_probe_cuda.cu:10-66: dst[i] = src[i] over two buffers,
// best of N).  One 16-byte load and store per thread over the whole buffer;
// reported as (read + write bytes) / kernel time.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "rtm_hip.h"

namespace {
__global__ __launch_bounds__(256) void k_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n4) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n4) dst[i] = src[i];
}
}  // namespace

extern "C" int rtm_hip_bandwidth_probe(int device, size_t bytes, int repeats, double* best_gbs) {
    if (!best_gbs || repeats < 1 || bytes < 4096) return RTM_HIP_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return RTM_HIP_ERR_DEVICE;
    const size_t n4 = bytes / 16;
    float4 *a = nullptr, *b = nullptr;
    if (hipMalloc(&a, n4 * 16) != hipSuccess) return RTM_HIP_ERR_MEMORY;
    if (hipMalloc(&b, n4 * 16) != hipSuccess) {
        (void)hipFree(a);
        return RTM_HIP_ERR_MEMORY;
    }
    const size_t blocks = (n4 + 255) / 256;
    if (blocks > 0x7fffffffULL) {
        (void)hipFree(a);
        (void)hipFree(b);
        return RTM_HIP_ERR_ARG;
    }
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    int rc = hipMemset(a, 0, n4 * 16) == hipSuccess ? RTM_HIP_OK : RTM_HIP_ERR_DEVICE;
    double best = 0.0;
    for (int r = 0; r < repeats + 1 && rc == RTM_HIP_OK; ++r) {
        (void)hipEventRecord(t0, nullptr);
        hipLaunchKernelGGL(k_copy, dim3((unsigned)blocks), dim3(256), 0, nullptr, a, b, n4);
        (void)hipEventRecord(t1, nullptr);
        if (hipEventSynchronize(t1) != hipSuccess) { rc = RTM_HIP_ERR_DEVICE; break; }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, t0, t1);
        if (r > 0 && ms > 0.0f) best = std::max(best, 2.0 * (double)(n4 * 16) / (ms * 1e-3) * 1e-9);
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    (void)hipFree(a);
    (void)hipFree(b);
    *best_gbs = best;
    return rc;
}
