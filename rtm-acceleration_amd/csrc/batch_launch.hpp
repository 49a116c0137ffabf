// batch_launch.hpp -- the batched step kernel (one launch per tick for
// every job of one stencil half-width; batch_kernels.hip) and its launch
// dispatch, templated on the snapshot width: batch_kernels_f32.hip and
// batch_kernels_s16.hip each build one half (jobs with FP32 and with 16-bit
// snapshots run in separate launches).
#pragma once

#include "batch_args.hpp"
#include "batch_kernels.hpp"
#include "march_tile.hpp"

namespace rtm_hip {
namespace {

using namespace device;

// The bookkeeping tick of a 16-bit job, between its phases (wave t of the
// job's tx_cols-column tiles): the stored forward steps among the last two have
// no step it+2 to settle their scale.  Step nt-2: its codes, from its field
// (still in its buffer), with the max of its slices; step nt-1: the slices of
// its per-tile maxima, which backward step 0 reduces to the max it images
// with.
template <int H>
__device__ __forceinline__ void bookkeeping_tile(const JobDesc& J, int t, int lane, int tx_cols) {
    constexpr int TZ = kTileZ - 2 * H;
    const int st = J.store, nt = J.nt;
    if (nt >= 2 && ((nt - 2) % st) == 0) {
        const int slot = (nt - 2) / st, p = slot & 1;
        const float m = parts_max(J.slice_max + p * kReduceWaves, J.part_counts + 2 + p, lane);
        const int k = snapshot_scale_exponent(m);
        if (t == 0 && lane == 0) gbl(J.snap_max)[slot] = m;
        const float* f = rotate(J, nt - 2);          // buffer of forward step nt-2's field
        uint16_t* out = J.snap16 + (ptrdiff_t)slot * J.n_int;
        const int tz = t % J.ntz, tx = t / J.ntz;
        const int z = tz * TZ + lane;
        const int jz = z - J.nb;
        const bool z_int = lane >= H && lane < kTileZ - H && z < J.nze - H && jz >= 0 && jz < J.nz;
        for (int r = 0; r < tx_cols; ++r) {
            const int x = H + tx * tx_cols + r, jx = x - J.nb;
            if (z_int && x < J.nxe - H && jx >= 0 && jx < J.nx)
                __builtin_nontemporal_store(encode_snapshot(gbl(f)[(ptrdiff_t)x * J.ldz + z], k, J.snap_fmt),
                                            gbl(out) + (ptrdiff_t)jx * J.nz + jz);
        }
    }
    if (((nt - 1) % st) == 0) {
        const int p = ((nt - 1) / st) & 1;
        reduce_slice(J.wave_max + (ptrdiff_t)p * J.parts_cap, J.part_counts + p, J.parts_cap,
                     J.slice_max + p * kReduceWaves, J.part_counts + 2 + p, t, J.ntiles, lane);
    }
}

// Tick `sub` of a run of ticks for every job of one group: the wave's job
// (wave-uniform, read-first-lane) and that job's tick arguments, read with
// independent scalar loads; then the same march/stream tile as the per-shot
// kernel.  LEAN: a tick in which no job of the group has snapshot work, run
// on the lean build of the march tile.
template <int H, int TX, Family F, bool S16, bool LEAN = false>
__global__ __launch_bounds__(kBlockThreads) void k_batch(const JobDesc* __restrict__ jobs,
                                                         const TickArgs* __restrict__ targs,
                                                         const int* __restrict__ tile_job, int ntiles, int sub) {
    const int wave = __builtin_amdgcn_readfirstlane((int)threadIdx.x / kTileZ);
    const int lane = (int)threadIdx.x & (kTileZ - 1);
    const int nblocks = (ntiles + kWaves - 1) / kWaves;
    const int g = xcd_tile((int)blockIdx.x, nblocks) * kWaves + wave;
    if (g >= ntiles) return;
    const int jid = __builtin_amdgcn_readfirstlane(tile_job[g]);
    const TickArgs& T = targs[jid * kRunTicks + sub];
    // the tick's arguments in one batch with its kind (request_args)
    asm volatile("" ::"s"(T.kind), "s"(T.tile_first));
    request_args(T.a);
    const int kind = T.kind;
    if (kind == kTickIdle) return;           // finished job: its tiles idle
    const int t = g - T.tile_first;
    const StepArgs& a = T.a;
    if (S16 && kind == kTickBookkeeping) {
        if (!a.dry) bookkeeping_tile<H>(jobs[jid], t, lane, TX);
        return;
    }
    if constexpr (F == Family::kStream) {
        if (kind == kTickForward) stream_tile<H, TX, Dir::Forward, S16>(a, a.coef, t, lane);
        else stream_tile<H, TX, Dir::Backward, S16>(a, a.coef, t, lane);
    } else {
        constexpr bool DPP = F == Family::kMarchDpp;
        if (kind == kTickForward) march_tile<H, TX, Dir::Forward, DPP, S16, LEAN>(a, a.coef, t, lane);
        else march_tile<H, TX, Dir::Backward, DPP, S16, LEAN>(a, a.coef, t, lane);
    }
}

template <int H, bool S16>
hipError_t launch_h(Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs, const int* tile_job,
                    int ntiles, int sub, size_t lds, hipStream_t s, bool lean) {
    const dim3 block(kBlockThreads);
    const dim3 grid((unsigned)((ntiles + kWaves - 1) / kWaves));
#define RTM_BATCH(TX, F) \
    hipLaunchKernelGGL((k_batch<H, TX, F, S16>), grid, block, lds, s, jobs, targs, tile_job, ntiles, sub)
#define RTM_BATCH_LEAN(TX) \
    hipLaunchKernelGGL((k_batch<H, TX, Family::kMarchDpp, S16, true>), grid, block, lds, s, jobs, targs, tile_job, \
                       ntiles, sub)
    if (lean && family == Family::kMarchDpp) {          // the DPP march family has lean builds
        switch (tile_x) {
            case 2: RTM_BATCH_LEAN(2); break;
            case 3: RTM_BATCH_LEAN(3); break;
            case 4: RTM_BATCH_LEAN(4); break;
            case 5: RTM_BATCH_LEAN(5); break;
            case 6: RTM_BATCH_LEAN(6); break;
            case 8: RTM_BATCH_LEAN(8); break;
            case 16: RTM_BATCH_LEAN(16); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
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
#undef RTM_BATCH_LEAN
    return hipGetLastError();
}


template <bool S16>
hipError_t launch_batch_step_t(int half, Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs,
                               const int* tile_job, int ntiles, int sub, int lds_bytes, hipStream_t stream,
                               bool lean) {
    const size_t lds = (size_t)lds_bytes;
    if (ntiles <= 0) return hipSuccess;
    switch (half) {
        case 1: return launch_h<1, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        case 2: return launch_h<2, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        case 3: return launch_h<3, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        case 4: return launch_h<4, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        case 5: return launch_h<5, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        case 6: return launch_h<6, S16>(family, tile_x, jobs, targs, tile_job, ntiles, sub, lds, stream, lean);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace
}  // namespace rtm_hip
