// Device post-processing of the stacked image (src/rtm/imaging.cpp:48-75).
#include "image_post.hpp"

namespace rtm_hip {
namespace {

constexpr int kBlock = 256;

// The reference's running maximum is std::max(smax, s), i.e. (smax < s) ? s : smax
// (src/rtm/imaging.cpp:51-52); a NaN is never taken and +0 is never replaced
// by -0, so any order of the same comparisons gives the same value.
__device__ inline float take_max(float m, float v) { return m < v ? v : m; }

__device__ float block_max(float m, float* red) {
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = take_max(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    return red[0];
}

// The FP64 stacks rounded once to float, as their download does.
__global__ void __launch_bounds__(kBlock)
k_round_stacks(const double* __restrict__ si, const double* __restrict__ ss, size_t n,
               float* __restrict__ fi, float* __restrict__ fs) {
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
        fi[i] = (float)si[i];
        fs[i] = (float)ss[i];
    }
}

// One partial maximum of S per block (grid-stride), kPostMaxParts blocks.
__global__ void __launch_bounds__(kBlock)
k_max_parts(const float* __restrict__ s, size_t n, float* __restrict__ parts) {
    __shared__ float red[kBlock];
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        m = take_max(m, s[i]);
    m = block_max(m, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = m;
}

// out = I / (S + eps), eps = eps_relative * max(S) + 1e-30f; every block
// reduces the partial maxima itself, so no launch is spent on the last step
// of the reduction.
__global__ void __launch_bounds__(kBlock)
k_illum_comp(const float* __restrict__ img, const float* __restrict__ s, size_t n,
             const float* __restrict__ parts, float eps_relative, float* __restrict__ out) {
    __shared__ float red[kBlock];
    float m = 0.0f;
    for (int j = threadIdx.x; j < kPostMaxParts; j += kBlock) m = take_max(m, parts[j]);
    const float smax = block_max(m, red);
    const float eps = eps_relative * smax + 1e-30f;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        out[i] = img[i] / (s[i] + eps);
}

// Second-order Laplacian, zero on the outer ring (src/rtm/imaging.cpp:63-75):
// idx2 * ((I[x+1] - 2 I) + I[x-1]) + idz2 * ((I[z+1] - 2 I) + I[z-1]).
__global__ void __launch_bounds__(kBlock)
k_laplacian(const float* __restrict__ in, int nx, int nz, float idx2, float idz2, float* __restrict__ out) {
    const int iz = blockIdx.x * kBlock + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= nz) return;
    const size_t i = (size_t)ix * nz + iz;
    float r = 0.0f;
    if (ix >= 1 && ix < nx - 1 && iz >= 1 && iz < nz - 1) {
        const float c = in[i];
        r = idx2 * (in[i + nz] - 2.0f * c + in[i - nz]) + idz2 * (in[i + 1] - 2.0f * c + in[i - 1]);
    }
    out[i] = r;
}

unsigned stride_blocks(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

}  // namespace

hipError_t launch_image_post(const float* image, const float* illum, const double* stack_i,
                             const double* stack_s, int nx, int nz, float idx2, float idz2,
                             int flags, float eps_relative, float* scratch, float* out,
                             hipStream_t stream) {
    const size_t n = (size_t)nx * nz;
    if (n == 0 || !(flags & 3) || nx > 65535) return hipErrorInvalidValue;
    float* parts = scratch;
    float* tmp = scratch + kPostMaxParts;
    if (stack_i) {
        float* fi = tmp + n;
        float* fs = tmp + 2 * n;
        hipLaunchKernelGGL(k_round_stacks, dim3(stride_blocks(n)), dim3(kBlock), 0, stream, stack_i, stack_s, n,
                           fi, fs);
        image = fi;
        illum = fs;
    }
    const float* cur = image;
    if (flags & 1) {
        float* dst = (flags & 2) ? tmp : out;
        hipLaunchKernelGGL(k_max_parts, dim3(kPostMaxParts), dim3(kBlock), 0, stream, illum, n, parts);
        hipLaunchKernelGGL(k_illum_comp, dim3(stride_blocks(n)), dim3(kBlock), 0, stream, image, illum, n, parts,
                           eps_relative, dst);
        cur = dst;
    }
    if (flags & 2) {
        hipLaunchKernelGGL(k_laplacian, dim3((unsigned)((nz + kBlock - 1) / kBlock), (unsigned)nx), dim3(kBlock), 0,
                           stream, cur, nx, nz, idx2, idz2, out);
    } else if (!(flags & 1)) {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rtm_hip
