// batch_kernels_s16.hip -- the batched step kernel for jobs with 16-bit (FP16/BF16)
// snapshots (batch_launch.hpp).
#include "batch_launch.hpp"

namespace rtm_hip {

hipError_t launch_batch_step_s16(int half, Family family, int tile_x, const JobDesc* jobs, const TickArgs* targs,
                                 const int* tile_job, int ntiles, int sub, int lds_bytes, hipStream_t stream,
                                 bool lean) {
    return launch_batch_step_t<true>(half, family, tile_x, jobs, targs, tile_job, ntiles, sub, lds_bytes, stream, lean);
}

}  // namespace rtm_hip
