// batch_kernels_merged.hip -- one launch per tick for the whole batch.
//
// The per-group launches (batch_launch.hpp) run jobs of one stencil
// half-width and snapshot width each; with several groups a tick is several
// launches, each with its own ramp and drain.  Here every job's tiles are in
// one launch: a wave reads its job's build with its tick arguments and runs,
// in a wave-uniform switch, the march tile built for that half-width and
// snapshot width, lean when the job has no snapshot work on this tick.  The
// kernel holds one tile body per (half-width, snapshot width, lean,
// direction); a wave only ever runs one, and the kernel's registers are the
// largest body's.
#include "batch_launch.hpp"

namespace rtm_hip {
namespace {

using namespace device;

template <int H, int TX, bool S16, bool LEAN>
__device__ __forceinline__ void merged_tile(const JobDesc& J, const StepArgs& a, int kind, int t, int lane) {
    if (S16 && kind == kTickBookkeeping) {
        if (!a.dry) bookkeeping_tile<H>(J, t, lane, TX);
        return;
    }
    if (kind == kTickForward) march_tile<H, TX, Dir::Forward, true, S16, LEAN>(a, a.coef, t, lane);
    else march_tile<H, TX, Dir::Backward, true, S16, LEAN>(a, a.coef, t, lane);
}

template <int TX>
__global__ __launch_bounds__(kBlockThreads) void k_batch_merged(const JobDesc* __restrict__ jobs,
                                                                const TickArgs* __restrict__ targs,
                                                                const int* __restrict__ tile_job, int ntiles, int sub) {
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (ntiles + kWaves - 1) / kWaves;
    const int g = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (g >= ntiles) return;
    const int jid = __builtin_amdgcn_readfirstlane(tile_job[g]);
    const TickArgs& T = targs[jid * kRunTicks + sub];
    // the tick's arguments in one batch with its kind and build (request_args)
    asm volatile("" ::"s"(T.kind), "s"(T.tile_first), "s"(T.build));
    request_args(T.a);
    const int kind = T.kind;
    if (kind == kTickIdle) return;           // finished job: its tiles idle
    const int t = g - T.tile_first;
    const StepArgs& a = T.a;
    const JobDesc& J = jobs[jid];
    // the job's build, uniform over the wave (TickArgs::build)
#define RTM_BUILDS(H)                                                                     \
    case 4 * H: merged_tile<H, TX, false, false>(J, a, kind, t, lane); break;             \
    case 4 * H + 1: merged_tile<H, TX, false, true>(J, a, kind, t, lane); break;          \
    case 4 * H + 2: merged_tile<H, TX, true, false>(J, a, kind, t, lane); break;          \
    case 4 * H + 3: merged_tile<H, TX, true, true>(J, a, kind, t, lane); break;
    switch (__builtin_amdgcn_readfirstlane(T.build)) {
        RTM_BUILDS(1)
        RTM_BUILDS(2)
        RTM_BUILDS(3)
        RTM_BUILDS(4)
        RTM_BUILDS(5)
        RTM_BUILDS(6)
        default: break;
    }
#undef RTM_BUILDS
}

}  // namespace

hipError_t launch_batch_merged(int tile_x, const JobDesc* jobs, const TickArgs* targs, const int* tile_job,
                               int ntiles, int sub, hipStream_t stream) {
    if (ntiles <= 0) return hipSuccess;
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves));
    switch (tile_x) {
        case 2: hipLaunchKernelGGL((k_batch_merged<2>), grid, block, 0, stream, jobs, targs, tile_job, ntiles, sub); break;
        case 3: hipLaunchKernelGGL((k_batch_merged<3>), grid, block, 0, stream, jobs, targs, tile_job, ntiles, sub); break;
        case 4: hipLaunchKernelGGL((k_batch_merged<4>), grid, block, 0, stream, jobs, targs, tile_job, ntiles, sub); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace rtm_hip
