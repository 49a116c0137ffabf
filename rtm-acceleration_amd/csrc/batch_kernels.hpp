// batch_kernels.hpp -- many shots, one launch per time step.
//
// Each job is one shot on its own grid (its own nx, nz, order, nt, store
// interval and tables).  The device holds one descriptor per job; a tick
// advances every active job by one step of its current phase (forward, then
// backward with imaging) with one launch per stencil half-width present, and
// a one-thread-per-job kernel then moves the job clocks and prepares every
// job's step arguments for the next tick.  The launch
// arguments never change between ticks, so a tick can be captured once in a
// hipGraph and replayed.
#pragma once

#include <hip/hip_runtime.h>

#include "step_kernels.hpp"

namespace rtm_hip {

struct JobDesc {
    // grid and schedule
    int nxe, nze, ldz, nb, nx, nz, half;
    int nt, store, snap_fmt;
    int ntz, ntx, ntiles;        // this job's tile grid (march mapping)
    int tile_first;              // its first tile in the group launch
    int bwd_base;                // buffer rotation base of the backward phase
    // clock, advanced on the device after every tick
    int phase;                   // 0 forward, 1 backward, 2 done
    int step;                    // steps completed in the phase
    Coeffs coef;
    long long trace_stride;
    long long n_int;
    const float* vdt2;
    const float* wx;
    const float* wz;
    const float* wavelet;
    const float* traces;
    float* fld[3];
    float* snap32;               // FP32 snapshots, or
    uint16_t* snap16;            // 16-bit codes +
    float* snap_max;             // per-snapshot max|p|
    float* image;
    float* illum;
    const int* tile_ptr[2];      // per-tile injection tables (forward, backward)
    const Entry* entries[2];
};

// One tick for the jobs of one half-width H: tile_job[g] names the job of
// global tile g, args[j] holds job j's step arguments at its clock (written
// by launch_batch_clock); family kMarch, kMarchDpp or kStream with the tile
// widths of the per-shot kernels.
hipError_t launch_batch_step(int half, Family family, int tile_x, const JobDesc* jobs, const StepArgs* args,
                             const int* tile_job, int ntiles, hipStream_t stream);

// FP64 stack update of a finished job: stack[(ix0+ix)*snz + iz0+iz] += image[ix*nz+iz]
// (and the same for illumination), for the job's nx x nz image.
hipError_t launch_stack_add(const float* image, const float* illum, int nx, int nz, double* stack_i,
                            double* stack_s, int stack_nz, int ix0, int iz0, hipStream_t stream);

// Job clocks and step arguments: advance != 0 moves every clock one step
// (at nt, to the next phase) first; then args[j] = the step arguments of job
// j at its clock, with the dry flag given (dry steps write only p^{n+1}: no
// image, illumination or snapshot update; used to time candidate kernels on
// live state).
hipError_t launch_batch_clock(JobDesc* jobs, StepArgs* args, int njobs, int advance, int dry, hipStream_t stream);

}  // namespace rtm_hip
