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

// Step arguments of job J at its current clock.
__device__ __forceinline__ StepArgs job_args(const JobDesc& J, int dry) {
    StepArgs a{};
    const int fwd = J.phase == 0;
    const int s = J.step;
    const int base = fwd ? 0 : J.bwd_base;
    a.p_prev = J.fld[(base + s) % 3];
    a.p_cur = J.fld[(base + s + 1) % 3];
    a.p_next = J.fld[(base + s + 2) % 3];
    a.zero_prev = s <= 1;
    a.zero_cur = s == 0;
    a.vdt2 = J.vdt2;
    a.wx = J.wx;
    a.wz = J.wz;
    a.coef = J.coef;
    a.nxe = J.nxe; a.nze = J.nze; a.ldz = J.ldz; a.nb = J.nb; a.nx = J.nx; a.nz = J.nz;
    a.ntz = J.ntz; a.ntx = J.ntx; a.ntiles = J.ntiles;
    const int it = fwd ? s : J.nt - 1 - s;
    a.it = it;
    a.tile_ptr = J.tile_ptr[fwd ? 0 : 1];
    a.entries = J.entries[fwd ? 0 : 1];
    a.wavelet = J.wavelet;
    a.traces = J.traces;
    a.trace_stride = J.trace_stride;
    a.image = J.image;
    a.illum = J.illum;
    a.dry = dry;
    a.snap_fmt = J.snap_fmt;
    const bool stored = (it % J.store) == 0;
    const long long slot = it / J.store;
    if (J.snap_fmt == (int)SnapFmt::F32) {
        if (stored) {
            if (fwd) a.snap_out = J.snap32 + slot * J.n_int;
            else a.snap_in = J.snap32 + slot * J.n_int;
        }
    } else if (fwd) {
        if (stored) a.snap_max = reinterpret_cast<unsigned*>(J.snap_max + slot);
        if (it >= 1 && ((it - 1) % J.store) == 0) {
            const long long prev = (it - 1) / J.store;
            a.quant_out = J.snap16 + prev * J.n_int;
            a.quant_max = J.snap_max + prev;
        }
    } else if (stored) {
        a.snap16_max = J.snap_max + slot;
        if (s == 0) a.pending_field = J.fld[J.bwd_base];   // last forward field
        else a.snap16_in = J.snap16 + slot * J.n_int;
    }
    return a;
}

template <int H, int TX, bool DPP>
__global__ __launch_bounds__(kBlockThreads) void k_batch(const JobDesc* __restrict__ jobs,
                                                         const int* __restrict__ tile_job, int ntiles, int dry) {
    const int wave = (int)threadIdx.x / kTileZ;
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (ntiles + kWaves - 1) / kWaves;
    const int g = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (g >= ntiles) return;
    const JobDesc& J = jobs[tile_job[g]];
    if (J.phase > 1) return;                 // finished job: its tiles idle
    const StepArgs a = job_args(J, dry);
    const int t = g - J.tile_first;
    if (J.phase == 0) march_tile<H, TX, Dir::Forward, DPP>(a, t, lane);
    else march_tile<H, TX, Dir::Backward, DPP>(a, t, lane);
}

__global__ __launch_bounds__(64) void k_advance(JobDesc* jobs, int njobs) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= njobs) return;
    JobDesc& J = jobs[j];
    if (J.phase > 1) return;
    const int s = J.step + 1;
    if (s == J.nt) {
        J.step = 0;
        J.phase = J.phase + 1;
    } else {
        J.step = s;
    }
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
hipError_t launch_h(int tile_x, bool dpp, const JobDesc* jobs, const int* tile_job, int ntiles, int dry,
                    hipStream_t s) {
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves));
#define RTM_BATCH(TX)                                                                                     \
    do {                                                                                                  \
        if (dpp) hipLaunchKernelGGL((k_batch<H, TX, true>), grid, block, 0, s, jobs, tile_job, ntiles, dry); \
        else hipLaunchKernelGGL((k_batch<H, TX, false>), grid, block, 0, s, jobs, tile_job, ntiles, dry);  \
    } while (0)
    switch (tile_x) {
        case 4: RTM_BATCH(4); break;
        case 8: RTM_BATCH(8); break;
        case 16: RTM_BATCH(16); break;
        default: return hipErrorInvalidValue;
    }
#undef RTM_BATCH
    return hipGetLastError();
}

}  // namespace

hipError_t launch_batch_step(int half, int tile_x, bool dpp, const JobDesc* jobs, const int* tile_job,
                             int ntiles, int dry, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    switch (half) {
        case 1: return launch_h<1>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
        case 2: return launch_h<2>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
        case 3: return launch_h<3>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
        case 4: return launch_h<4>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
        case 5: return launch_h<5>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
        case 6: return launch_h<6>(tile_x, dpp, jobs, tile_job, ntiles, dry, stream);
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

hipError_t launch_batch_advance(JobDesc* jobs, int njobs, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, stream, jobs, njobs);
    return hipGetLastError();
}

}  // namespace rtm_hip
