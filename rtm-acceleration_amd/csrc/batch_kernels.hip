// batch_kernels.hip -- the batched step: every wave finds its job and tile,
// rebuilds that job's step arguments from the descriptor and runs the same
// march tile as the per-shot engine (march_tile.hpp), so a job's image in a
// batch is bit-identical to its image alone.
//
// Phase boundaries need no reset pass: in the first step of a phase p^n and
// p^{n-1} are read as zero, and in the second p^{n-1} is (their buffers still
// hold the other phase's data); the reference zeroes all three buffers
// instead (src/rtm/backward.cpp:46-48), which gives the same arithmetic.  The
// backward phase rotates from the buffer that holds the forward phase's last
// field, so a 16-bit snapshot of that last step can be taken at backward step
// 0 straight from it.
#include "batch_kernels.hpp"

#include "march_tile.hpp"

namespace rtm_hip {

namespace {

using namespace device;

__device__ __forceinline__ float* rotate(const JobDesc& J, int r) {
    return r == 0 ? J.fld[0] : (r == 1 ? J.fld[1] : J.fld[2]);
}

// Step arguments of job J after `total` of its 2 nt steps (forward steps
// 0..nt-1, then backward steps 0..nt-1).  The stencil coefficients are read
// from the descriptor where they are used (an array copied into a local
// struct would be kept in scratch memory).
__device__ __forceinline__ StepArgs job_args(const JobDesc& J, int total, int dry) {
    StepArgs a{};
    const int fwd = total < J.nt;
    const int s = fwd ? total : total - J.nt;
    const int base = fwd ? 0 : J.bwd_base;
    a.p_prev = rotate(J, (base + s) % 3);
    a.p_cur = rotate(J, (base + s + 1) % 3);
    a.p_next = rotate(J, (base + s + 2) % 3);
    a.zero_prev = s <= 1;
    a.zero_cur = s == 0;
    a.vdt2 = J.vdt2;
    a.wx = J.wx;
    a.wz = J.wz;
    a.nxe = J.nxe; a.nze = J.nze; a.ldz = J.ldz; a.nb = J.nb; a.nx = J.nx; a.nz = J.nz;
    a.ntz = J.ntz; a.ntx = J.ntx; a.ntiles = J.ntiles;
    const int it = fwd ? s : J.nt - 1 - s;
    a.it = it;
    a.tile_ptr = fwd ? J.tile_ptr[0] : J.tile_ptr[1];
    a.entries = fwd ? J.entries[0] : J.entries[1];
    a.wavelet = J.wavelet;
    a.traces = J.traces;
    a.trace_stride = J.trace_stride;
    a.image = J.image;
    a.illum = J.illum;
    a.dry = dry;
    a.snap_fmt = J.snap_fmt;
    // Every field is assigned once, unconditionally: stores of different
    // fields in different branches can be merged into one store at a
    // computed offset, which would keep the whole struct in scratch memory.
    const bool stored = (it % J.store) == 0;
    const long long slot = it / J.store;
    const bool wide = J.snap_fmt == (int)SnapFmt::F32;
    const bool quant = !wide && fwd && it >= 1 && ((it - 1) % J.store) == 0;
    const long long prev = quant ? (it - 1) / J.store : 0;
    a.snap_out = (wide && fwd && stored) ? J.snap32 + slot * J.n_int : nullptr;
    a.snap_in = (wide && !fwd && stored) ? J.snap32 + slot * J.n_int : nullptr;
    a.wave_max = (!wide && fwd && stored) ? J.wave_max : nullptr;
    a.quant_out = quant ? J.snap16 + prev * J.n_int : nullptr;
    a.quant_max = quant ? J.snap_max + prev : nullptr;
    a.snap16_max = (!wide && !fwd && stored) ? J.snap_max + slot : nullptr;
    // the last forward field, at backward step 0
    a.pending_field = (!wide && !fwd && stored && s == 0) ? rotate(J, J.bwd_base) : nullptr;
    a.snap16_in = (!wide && !fwd && stored && s != 0) ? J.snap16 + slot * J.n_int : nullptr;
    return a;
}

// Tick `sub` of a run of ticks for every job of half-width H: job J is at
// step J.done + sub.  The job and tile of a wave are wave-uniform
// (read-first-lane), so the descriptor is read with scalar loads and the
// step arguments are built in scalar registers, as a kernel argument would
// be.
template <int H, int TX, Family F>
__global__ __launch_bounds__(kBlockThreads) void k_batch(const JobDesc* __restrict__ jobs,
                                                         const int* __restrict__ tile_job, int ntiles, int sub,
                                                         int dry) {
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (ntiles + kWaves - 1) / kWaves;
    const int g = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (g >= ntiles) return;
    const int jid = __builtin_amdgcn_readfirstlane(tile_job[g]);
    const JobDesc& J = jobs[jid];
    const int total = J.done + sub;
    if (total >= 2 * J.nt) return;           // finished job: its tiles idle
    const StepArgs a = job_args(J, total, dry);
    const int t = g - J.tile_first;
    if constexpr (F == Family::kStream) {
        if (total < J.nt) stream_tile<H, TX, Dir::Forward>(a, J.coef, t, lane);
        else stream_tile<H, TX, Dir::Backward>(a, J.coef, t, lane);
    } else {
        constexpr bool DPP = F == Family::kMarchDpp;
        if (total < J.nt) march_tile<H, TX, Dir::Forward, DPP>(a, J.coef, t, lane);
        else march_tile<H, TX, Dir::Backward, DPP>(a, J.coef, t, lane);
    }
}

__global__ __launch_bounds__(256) void k_batch_reduce(const JobDesc* __restrict__ jobs, int sub) {
    const JobDesc& J = jobs[blockIdx.x];
    const int it = J.done + sub;             // forward step just done, if any
    if (J.snap_fmt == (int)SnapFmt::F32 || it >= J.nt || (it % J.store) != 0) return;   // whole block
    float m = 0.0f;
    for (int i = (int)threadIdx.x; i < J.ntiles; i += (int)blockDim.x) m = fmaxf(m, J.wave_max[i]);
    m = wave_max(m);
    __shared__ float wm[256 / kTileZ];
    if ((threadIdx.x & (kTileZ - 1)) == 0) wm[threadIdx.x / kTileZ] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        float r = 0.0f;
        for (int k = 0; k < 256 / kTileZ; ++k) r = fmaxf(r, wm[k]);
        J.snap_max[it / J.store] = r;
    }
}

// Job clocks: every job has done `ticks` more steps.
__global__ __launch_bounds__(64) void k_advance(JobDesc* jobs, int njobs, int ticks) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j < njobs) jobs[j].done += ticks;
}

__global__ __launch_bounds__(256) void k_stack_add(const float* image, const float* illum, int nx, int nz,
                                                   double* si, double* ss, int snz, int ix0, int iz0) {
    const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= (long long)nx * nz) return;
    const int ix = (int)(j / nz), iz = (int)(j % nz);
    const long long k = (long long)(ix0 + ix) * snz + (iz0 + iz);
    si[k] += (double)image[j];
    ss[k] += (double)illum[j];
}

template <int H>
hipError_t launch_h(Family family, int tile_x, const JobDesc* jobs, const int* tile_job, int ntiles, int sub,
                    int dry, size_t lds, hipStream_t s) {
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves));
#define RTM_BATCH(TX, F) \
    hipLaunchKernelGGL((k_batch<H, TX, F>), grid, block, lds, s, jobs, tile_job, ntiles, sub, dry)
    if (family == Family::kStream) {
        switch (tile_x) {
            case 8: RTM_BATCH(8, Family::kStream); break;
            case 16: RTM_BATCH(16, Family::kStream); break;
            case 32: RTM_BATCH(32, Family::kStream); break;
            default: return hipErrorInvalidValue;
        }
    } else if (family == Family::kMarchDpp) {
        switch (tile_x) {
            case 2: RTM_BATCH(2, Family::kMarchDpp); break;
            case 3: RTM_BATCH(3, Family::kMarchDpp); break;
            case 4: RTM_BATCH(4, Family::kMarchDpp); break;
            case 5: RTM_BATCH(5, Family::kMarchDpp); break;
            case 6: RTM_BATCH(6, Family::kMarchDpp); break;
            case 8: RTM_BATCH(8, Family::kMarchDpp); break;
            case 16: RTM_BATCH(16, Family::kMarchDpp); break;
            default: return hipErrorInvalidValue;
        }
    } else if (family == Family::kMarch) {
        switch (tile_x) {
            case 4: RTM_BATCH(4, Family::kMarch); break;
            case 8: RTM_BATCH(8, Family::kMarch); break;
            case 16: RTM_BATCH(16, Family::kMarch); break;
            default: return hipErrorInvalidValue;
        }
    } else {
        return hipErrorInvalidValue;
    }
#undef RTM_BATCH
    return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_step(int half, Family family, int tile_x, const JobDesc* jobs, const int* tile_job,
                             int ntiles, int sub, int dry, int lds_bytes, hipStream_t stream) {
    const size_t lds = (size_t)lds_bytes;
    if (ntiles <= 0) return hipSuccess;
    switch (half) {
        case 1: return launch_h<1>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        case 2: return launch_h<2>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        case 3: return launch_h<3>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        case 4: return launch_h<4>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        case 5: return launch_h<5>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        case 6: return launch_h<6>(family, tile_x, jobs, tile_job, ntiles, sub, dry, lds, stream);
        default: return hipErrorInvalidValue;
    }
}

hipError_t launch_stack_add(const float* image, const float* illum, int nx, int nz, double* stack_i,
                            double* stack_s, int stack_nz, int ix0, int iz0, hipStream_t stream) {
    const long long n = (long long)nx * nz;
    hipLaunchKernelGGL(k_stack_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, image, illum, nx, nz,
                       stack_i, stack_s, stack_nz, ix0, iz0);
    return hipGetLastError();
}

hipError_t launch_batch_reduce(JobDesc* jobs, int njobs, int sub, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_reduce, dim3((unsigned)njobs), dim3(256), 0, stream, jobs, sub);
    return hipGetLastError();
}

hipError_t launch_batch_advance(JobDesc* jobs, int njobs, int ticks, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, stream, jobs, njobs, ticks);
    return hipGetLastError();
}

}  // namespace rtm_hip
