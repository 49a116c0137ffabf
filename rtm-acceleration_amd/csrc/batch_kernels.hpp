// batch_kernels.hpp -- many shots, one launch per time step.
//
// Each job is one shot on its own grid (its own nx, nz, order, nt, store
// interval and tables).  The device holds one descriptor per job with the
// number of steps it has done; a tick advances every active job by one step
// of its current phase (forward, then backward with imaging) with one launch
// per stencil half-width present.  A run of k ticks passes each launch its
// offset 0..k-1 within the run and ends with one launch that adds k to every
// job's count, so the launches of a run never change and a run can be
// captured once in a hipGraph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "step_kernels.hpp"

namespace rtm_hip {

struct JobDesc {
    // grid and schedule
    int nxe, nze, ldz, nb, nx, nz, half;
    int nt, store, snap_fmt;
    int ntz, ntx, ntiles;        // this job's tile grid (march mapping)
    unsigned ntz_magic;          // tile_div_magic(ntz)
    int tile_first;              // its first tile in the group launch
    int bwd_base;                // rotation parity of the backward phase's first output
    // clock, advanced on the device after every run of ticks
    int done;                    // steps done of `steps`
    int steps;                   // forward nt, then (16-bit snapshots) one
                                 // bookkeeping tick, then backward nt
    Coeffs coef;
    long long n_int;
    const float* vdt2;
    const float* wx;
    const float* wz;
    const float* wavelet;
    float* amps;                 // receiver amplitudes in backward-entry order (StepArgs::amps)
    int amp_count;
    float* fld[3];               // p^{n-1} (overwritten in place by p^{n+1}), p^n, and the
                                 // output of dry (timing) steps, which leave the state intact
    float* hold[2];              // fields held for paired imaging (pair_step)
    int pair_first;              // pair_first() of the backward phase
    const float* zero;           // a field of zeros (the model's), read as p^{n-1} / p^n on a phase's first steps
    float* snap32;               // FP32 snapshots, or
    uint16_t* snap16;            // 16-bit codes +
    float* snap_max;             // per-snapshot max|p|
    // scale bookkeeping of 16-bit snapshots (StepArgs), double-buffered by
    // snapshot parity: per-tile maxima [2][parts_cap], their slice maxima
    // [2][kReduceWaves], counts {tiles[2], slices[2]}
    float* wave_max;
    float* slice_max;
    int* part_counts;
    int parts_cap;
    float* image;
    float* illum;
    const int* tile_ptr[2];      // per-tile injection tables (forward, backward)
    const Entry* entries[2];
};

// Ticks per captured run; the tick-argument table holds this many entries
// per job.
constexpr int kRunTicks = 32;

// One job's arguments for one tick of a run, resolved once per run by
// launch_tick_args: the step arguments (buffers, slots, flags of that step)
// and what the tick is for the job.  The step kernel reads them with scalar
// loads that do not depend on each other, instead of resolving them from
// the descriptor in every wave (a chain of dependent scalar loads and integer
// divisions before the first vector load).
enum TickKind : int { kTickForward = 0, kTickBackward = 1, kTickBookkeeping = 2, kTickIdle = 3 };
struct TickArgs {
    StepArgs a;
    int kind;                    // TickKind
    int tile_first;              // the job's first tile in its group launch
    int build;                   // the job's tile build in the merged launch: 4 * half-width
                                 // + 2 (16-bit snapshots) + 1 (lean: no snapshot work this tick);
                                 // from the descriptor, as a bookkeeping tick's step arguments
                                 // are empty
};

// Tick arguments of ticks 0..nsub-1 of the coming run for every job:
// targs[job * kRunTicks + sub] (nsub <= kRunTicks).  dry != 0 marks them as
// timing launches (StepArgs::dry).
hipError_t launch_tick_args(const JobDesc* jobs, int njobs, TickArgs* targs, int nsub, int dry,
                            hipStream_t stream);

// Tick `sub` of a run for the jobs of one half-width H: tile_job[g] names
// the job of global tile g, and job j computes the step described by
// targs[j * kRunTicks + sub] (launch_tick_args); family kMarch, kMarchDpp or
// kStream with the tile widths of the per-shot kernels.  Tick arguments made
// with dry != 0 write only p^{n+1}: no image, illumination or snapshot
// update (used to time candidate kernels on live state).  lds_bytes of
// dynamic LDS per workgroup cap the workgroups resident on a CU.
// snap16 selects the build for jobs with 16-bit snapshots (every job of
// the launch must have the same snapshot width).
// lean: no job of the group has snapshot work on this tick (groups of the
// DPP march family then run a build without the snapshot epilogue).
hipError_t launch_batch_step(int half, Family family, int tile_x, bool snap16, const JobDesc* jobs,
                             const TickArgs* targs, const int* tile_job, int ntiles, int sub, int lds_bytes,
                             hipStream_t stream, bool lean = false);
hipError_t launch_batch_step_f32(int half, Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs,
                                 const int* tile_job, int ntiles, int sub, int lds_bytes, hipStream_t stream,
                                 bool lean);
hipError_t launch_batch_step_s16(int half, Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs,
                                 const int* tile_job, int ntiles, int sub, int lds_bytes, hipStream_t stream,
                                 bool lean);

// Tick `sub` of a run for every job of the batch in one launch, whatever
// its stencil half-width or snapshot width: tile_job covers all jobs, each
// with tile_x-column DPP march tiles, and each wave runs the march tile
// built for its job's half-width and snapshot width, in its lean build when
// the job has no snapshot work on this tick (TickArgs::build).  tile_x is
// one of merged_tile_widths().
hipError_t launch_batch_merged(int tile_x, const JobDesc* jobs, const TickArgs* targs, const int* tile_job,
                               int ntiles, int sub, hipStream_t stream);
inline std::vector<int> merged_tile_widths() { return {2, 3, 4}; }

// FP64 stack update of a finished job: stack[(ix0+ix)*snz + iz0+iz] += image[ix*nz+iz]
// (and the same for illumination), for the job's nx x nz image.
hipError_t launch_stack_add(const float* image, const float* illum, int nx, int nz, double* stack_i,
                            double* stack_s, int stack_nz, int ix0, int iz0, hipStream_t stream);

// Job clocks: every job has done `ticks` more steps.
hipError_t launch_batch_advance(JobDesc* jobs, int njobs, int ticks, hipStream_t stream);

// The bookkeeping tick of 16-bit job jobs[jid] alone, over its ntiles
// tile_x-column tiles (a job that runs on the per-shot step kernels).
hipError_t launch_bookkeeping(const JobDesc* jobs, int jid, int half, int tile_x, int ntiles, hipStream_t stream);

}  // namespace rtm_hip
