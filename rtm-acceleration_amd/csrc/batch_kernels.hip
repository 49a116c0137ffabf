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

// Step arguments of job J at its current clock.
__device__ __forceinline__ StepArgs job_args(const JobDesc& J, int dry) {
    StepArgs a{};
    const int fwd = J.phase == 0;
    const int s = J.step;
    const int base = fwd ? 0 : J.bwd_base;
    a.p_prev = rotate(J, (base + s) % 3);
    a.p_cur = rotate(J, (base + s + 1) % 3);
    a.p_next = rotate(J, (base + s + 2) % 3);
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
        if (s == 0) a.pending_field = rotate(J, J.bwd_base);   // last forward field
        else a.snap16_in = J.snap16 + slot * J.n_int;
    }
    return a;
}

// One tick of every job of half-width H.  The job and tile of a wave are
// wave-uniform (read-first-lane), so the job's step arguments, prepared by
// k_clock, are read with scalar loads as a kernel argument would be.
template <int H, int TX, Family F>
__global__ __launch_bounds__(kBlockThreads) void k_batch(const JobDesc* __restrict__ jobs,
                                                         const StepArgs* __restrict__ args,
                                                         const int* __restrict__ tile_job, int ntiles) {
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (ntiles + kWaves - 1) / kWaves;
    const int g = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (g >= ntiles) return;
    const int jid = __builtin_amdgcn_readfirstlane(tile_job[g]);
    const int phase = jobs[jid].phase;
    if (phase > 1) return;                   // finished job: its tiles idle
    const int t = g - jobs[jid].tile_first;
    if constexpr (F == Family::kStream) {
        if (phase == 0) stream_tile<H, TX, Dir::Forward>(args[jid], t, lane);
        else stream_tile<H, TX, Dir::Backward>(args[jid], t, lane);
    } else {
        constexpr bool DPP = F == Family::kMarchDpp;
        if (phase == 0) march_tile<H, TX, Dir::Forward, DPP>(args[jid], t, lane);
        else march_tile<H, TX, Dir::Backward, DPP>(args[jid], t, lane);
    }
}

// Job clocks (advance != 0: one step; at nt, the next phase) and the step
// arguments of every job at its clock.
__global__ __launch_bounds__(64) void k_clock(JobDesc* jobs, StepArgs* args, int njobs, int advance, int dry) {
    const int j = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (j >= njobs) return;
    JobDesc& J = jobs[j];
    if (J.phase > 1) return;
    if (advance) {
        const int s = J.step + 1;
        if (s == J.nt) {
            J.step = 0;
            J.phase = J.phase + 1;
        } else {
            J.step = s;
        }
        if (J.phase > 1) return;
    }
    args[j] = job_args(J, dry);
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
hipError_t launch_h(Family family, int tile_x, const JobDesc* jobs, const StepArgs* args, const int* tile_job,
                    int ntiles, hipStream_t s) {
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves));
#define RTM_BATCH(TX, F) hipLaunchKernelGGL((k_batch<H, TX, F>), grid, block, 0, s, jobs, args, tile_job, ntiles)
    if (family == Family::kStream) {
        switch (tile_x) {
            case 8: RTM_BATCH(8, Family::kStream); break;
            case 16: RTM_BATCH(16, Family::kStream); break;
            case 32: RTM_BATCH(32, Family::kStream); break;
            default: return hipErrorInvalidValue;
        }
    } else if (family == Family::kMarchDpp) {
        switch (tile_x) {
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

hipError_t launch_batch_step(int half, Family family, int tile_x, const JobDesc* jobs, const StepArgs* args,
                             const int* tile_job, int ntiles, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    switch (half) {
        case 1: return launch_h<1>(family, tile_x, jobs, args, tile_job, ntiles, stream);
        case 2: return launch_h<2>(family, tile_x, jobs, args, tile_job, ntiles, stream);
        case 3: return launch_h<3>(family, tile_x, jobs, args, tile_job, ntiles, stream);
        case 4: return launch_h<4>(family, tile_x, jobs, args, tile_job, ntiles, stream);
        case 5: return launch_h<5>(family, tile_x, jobs, args, tile_job, ntiles, stream);
        case 6: return launch_h<6>(family, tile_x, jobs, args, tile_job, ntiles, stream);
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

hipError_t launch_batch_clock(JobDesc* jobs, StepArgs* args, int njobs, int advance, int dry, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_clock, dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, stream, jobs, args, njobs,
                       advance, dry);
    return hipGetLastError();
}

}  // namespace rtm_hip
