// batch_kernels_f32.hip -- the batched step kernel for jobs with FP32
// snapshots (batch_launch.hpp).
#include "batch_launch.hpp"

namespace rtm_hip {

hipError_t launch_batch_step_f32(int half, Family family, int tile_x, const JobDesc* jobs, const int* tile_job,
                                 int ntiles, int sub, int dry, int lds_bytes, hipStream_t stream) {
    return launch_batch_step_t<false>(half, family, tile_x, jobs, tile_job, ntiles, sub, dry, lds_bytes, stream);
}

}  // namespace rtm_hip
