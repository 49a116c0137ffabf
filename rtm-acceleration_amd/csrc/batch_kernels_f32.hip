// batch_kernels_f32.hip -- the batched step kernel for jobs with FP32
// snapshots (batch_launch.hpp).
#include "batch_launch.hpp"

namespace rtm_hip {

hipError_t launch_batch_step_f32(int half, Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs,
                                 const int* tile_job, int ntiles, int sub, int lds_bytes, hipStream_t stream,
                                 bool lean) {
    return launch_batch_step_t<false>(half, family, tile_x, jobs, targs, tile_job, ntiles, sub, lds_bytes, stream, lean);
}

}  // namespace rtm_hip
