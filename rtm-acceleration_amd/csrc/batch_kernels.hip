// batch_kernels.hip -- the batched step: every wave finds its job and tile,
// rebuilds that job's step arguments from the descriptor and runs the same
// march tile as the per-shot engine (march_tile.hpp), so a job's image in a
// batch is bit-identical to its image alone.
//
// Phase boundaries need no reset pass: in the first step of a phase p^n and
// p^{n-1} are read as zero, and in the second p^{n-1} is (their buffers still
// hold the other phase's data); the reference zeroes all three buffers
// instead (src/rtm/backward.cpp:46-48), which gives the same arithmetic.  A
// job's two field buffers alternate, each step writing over its own p^{n-1}
// (batch_args.hpp); the backward phase starts in the buffer that does not
// hold the forward phase's last field, so a 16-bit snapshot of that last
// step can be taken at backward step 0 straight from it.
//
// The step kernel itself is in batch_launch.hpp, built once for jobs with
// FP32 snapshots and once for jobs with 16-bit ones; this file holds the
// clock and stack kernels and picks the build.
#include "batch_kernels.hpp"

#include "batch_launch.hpp"

namespace rtm_hip {

namespace {

// Tick arguments of ticks 0..nsub-1 of the coming run, one thread per job
// and tick.
__global__ __launch_bounds__(64) void k_tick_args(const JobDesc* jobs, int njobs, TickArgs* targs, int nsub,
                                                  int dry) {
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= njobs * nsub) return;
    const int jid = i / nsub, sub = i % nsub;
    const JobDesc& J = jobs[jid];
    const int total = J.done + sub;
    const bool s16 = J.snap_fmt != (int)SnapFmt::F32;
    TickArgs& T = targs[jid * kRunTicks + sub];
    T.kind = total >= J.steps ? kTickIdle
             : total < J.nt  ? kTickForward
             : (s16 && total == J.nt) ? kTickBookkeeping
                                      : kTickBackward;
    T.tile_first = J.tile_first;
    bool lean = false;
    if (T.kind == kTickForward || T.kind == kTickBackward) {
        T.a = s16 ? job_args<true>(J, total, dry) : job_args<false>(J, total, dry);
        lean = step_is_lean(T.kind == kTickForward ? Dir::Forward : Dir::Backward, T.a, s16);
    } else {                                 // bookkeeping reads the descriptor; idle nothing
        T.a = StepArgs{};
        T.a.dry = dry;
    }
    T.build = 4 * J.half + (s16 ? 2 : 0) + (lean ? 1 : 0);
}

// The bookkeeping tick of job jid on its own (a job running on the per-shot
// step kernels): one wave per tile of its tile_x-column tiles.
template <int H>
__global__ __launch_bounds__(kBlockThreads) void k_bookkeeping(const JobDesc* jobs, int jid, int tile_x) {
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const JobDesc& J = jobs[jid];
    const int t = (int)blockIdx.x * kWaves + wave;
    if (t >= J.ntiles) return;
    bookkeeping_tile<H>(J, t, lane, tile_x);
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

}  // namespace

hipError_t launch_batch_step(int half, Family family, int tile_x, bool snap16, const JobDesc* jobs,
                             const TickArgs* targs, const int* tile_job, int ntiles, int sub, int lds_bytes,
                             hipStream_t stream, bool lean) {
    return snap16 ? launch_batch_step_s16(half, family, tile_x, jobs, targs, tile_job, ntiles, sub, lds_bytes, stream,
                                          lean)
                  : launch_batch_step_f32(half, family, tile_x, jobs, targs, tile_job, ntiles, sub, lds_bytes, stream,
                                          lean);
}

hipError_t launch_bookkeeping(const JobDesc* jobs, int jid, int half, int tile_x, int ntiles, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves)), block(kBlockThreads);
    switch (half) {
        case 1: hipLaunchKernelGGL(k_bookkeeping<1>, grid, block, 0, stream, jobs, jid, tile_x); break;
        case 2: hipLaunchKernelGGL(k_bookkeeping<2>, grid, block, 0, stream, jobs, jid, tile_x); break;
        case 3: hipLaunchKernelGGL(k_bookkeeping<3>, grid, block, 0, stream, jobs, jid, tile_x); break;
        case 4: hipLaunchKernelGGL(k_bookkeeping<4>, grid, block, 0, stream, jobs, jid, tile_x); break;
        case 5: hipLaunchKernelGGL(k_bookkeeping<5>, grid, block, 0, stream, jobs, jid, tile_x); break;
        case 6: hipLaunchKernelGGL(k_bookkeeping<6>, grid, block, 0, stream, jobs, jid, tile_x); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_tick_args(const JobDesc* jobs, int njobs, TickArgs* targs, int nsub, int dry,
                            hipStream_t stream) {
    const int n = njobs * nsub;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_tick_args, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, stream, jobs, njobs, targs, nsub,
                       dry);
    return hipGetLastError();
}

hipError_t launch_stack_add(const float* image, const float* illum, int nx, int nz, double* stack_i,
                            double* stack_s, int stack_nz, int ix0, int iz0, hipStream_t stream) {
    const long long n = (long long)nx * nz;
    hipLaunchKernelGGL(k_stack_add, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, image, illum, nx, nz,
                       stack_i, stack_s, stack_nz, ix0, iz0);
    return hipGetLastError();
}

hipError_t launch_batch_advance(JobDesc* jobs, int njobs, int ticks, hipStream_t stream) {
    if (njobs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_advance, dim3((unsigned)((njobs + 63) / 64)), dim3(64), 0, stream, jobs, njobs, ticks);
    return hipGetLastError();
}

}  // namespace rtm_hip
