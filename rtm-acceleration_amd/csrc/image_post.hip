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

// A source value as float: the FP32 image, or the FP64 stack rounded once to
// float, as the stack's download rounds it.
__device__ __forceinline__ float src(const float* p, size_t i) { return p[i]; }
__device__ __forceinline__ float src(const double* p, size_t i) { return (float)p[i]; }

// One partial maximum of S per block (grid-stride), kPostMaxParts blocks.
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_max_parts(const T* __restrict__ s, size_t n, float* __restrict__ parts) {
    __shared__ float red[kBlock];
    float m = 0.0f;
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        m = take_max(m, src(s, i));
    m = block_max(m, red);
    if (threadIdx.x == 0) parts[blockIdx.x] = m;
}

// eps = eps_relative * max(S) + 1e-30f from the partial maxima, in every
// block (no launch is spent on the last step of the reduction).
__device__ float compensation_eps(const float* __restrict__ parts, float eps_relative, float* red) {
    float m = 0.0f;
    for (int j = threadIdx.x; j < kPostMaxParts; j += kBlock) m = take_max(m, parts[j]);
    return eps_relative * block_max(m, red) + 1e-30f;
}

// The compensated value of cell i, I / (S + eps).
template <typename T>
__device__ __forceinline__ float compensated(const T* img, const T* s, size_t i, float eps) {
    return src(img, i) / (src(s, i) + eps);
}

// out = I / (S + eps).
template <typename T>
__global__ void __launch_bounds__(kBlock)
k_illum_comp(const T* __restrict__ img, const T* __restrict__ s, size_t n, const float* __restrict__ parts,
             float eps_relative, float* __restrict__ out) {
    __shared__ float red[kBlock];
    const float eps = compensation_eps(parts, eps_relative, red);
    for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
        out[i] = compensated(img, s, i, eps);
}

// Second-order Laplacian, zero on the outer ring (src/rtm/imaging.cpp:63-75):
// idx2 * ((I[x+1] - 2 I) + I[x-1]) + idz2 * ((I[z+1] - 2 I) + I[z-1]).  COMP:
// of the compensated image, each of the five values compensated where it is
// read (the same float expression per cell), so no intermediate image is
// stored.
template <typename T, bool COMP>
__global__ void __launch_bounds__(kBlock)
k_laplacian(const T* __restrict__ img, const T* __restrict__ s, const float* __restrict__ parts,
            float eps_relative, int nx, int nz, float idx2, float idz2, float* __restrict__ out) {
    __shared__ float red[kBlock];
    float eps = 0.0f;
    if constexpr (COMP) eps = compensation_eps(parts, eps_relative, red);
    const int iz = blockIdx.x * kBlock + threadIdx.x;
    const int ix = blockIdx.y;
    if (iz >= nz) return;
    const size_t i = (size_t)ix * nz + iz;
    auto in = [&](size_t k) { return COMP ? compensated(img, s, k, eps) : src(img, k); };
    float r = 0.0f;
    if (ix >= 1 && ix < nx - 1 && iz >= 1 && iz < nz - 1) {
        const float c = in(i);
        r = idx2 * (in(i + nz) - 2.0f * c + in(i - nz)) + idz2 * (in(i + 1) - 2.0f * c + in(i - 1));
    }
    out[i] = r;
}

unsigned stride_blocks(size_t n) {
    const size_t b = (n + kBlock - 1) / kBlock;
    return (unsigned)(b < 4096 ? (b ? b : 1) : 4096);
}

template <typename T>
hipError_t launch_post(const T* image, const T* illum, int nx, int nz, float idx2, float idz2, int flags,
                       float eps_relative, float* parts, float* out, hipStream_t stream) {
    const size_t n = (size_t)nx * nz;
    if (flags & 1)
        hipLaunchKernelGGL((k_max_parts<T>), dim3(kPostMaxParts), dim3(kBlock), 0, stream, illum, n, parts);
    if (flags == 1) {
        hipLaunchKernelGGL((k_illum_comp<T>), dim3(stride_blocks(n)), dim3(kBlock), 0, stream, image, illum, n,
                           parts, eps_relative, out);
    } else {
        const dim3 grid((unsigned)((nz + kBlock - 1) / kBlock), (unsigned)nx);
        if (flags & 1)
            hipLaunchKernelGGL((k_laplacian<T, true>), grid, dim3(kBlock), 0, stream, image, illum, parts,
                               eps_relative, nx, nz, idx2, idz2, out);
        else
            hipLaunchKernelGGL((k_laplacian<T, false>), grid, dim3(kBlock), 0, stream, image, illum, parts,
                               eps_relative, nx, nz, idx2, idz2, out);
    }
    return hipGetLastError();
}

}  // namespace

hipError_t launch_image_post(const float* image, const float* illum, const double* stack_i,
                             const double* stack_s, int nx, int nz, float idx2, float idz2,
                             int flags, float eps_relative, float* scratch, float* out,
                             hipStream_t stream) {
    const size_t n = (size_t)nx * nz;
    if (n == 0 || flags < 1 || flags > 3 || nx > 65535) return hipErrorInvalidValue;
    return stack_i ? launch_post(stack_i, stack_s, nx, nz, idx2, idz2, flags, eps_relative, scratch, out, stream)
                   : launch_post(image, illum, nx, nz, idx2, idz2, flags, eps_relative, scratch, out, stream);
}

}  // namespace rtm_hip
