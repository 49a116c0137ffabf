// step_kernels_f32.hip -- the fused step kernels for FP32 snapshots
// (step_launch.hpp).
#include "step_launch.hpp"

namespace rtm_hip {

hipError_t launch_step_f32(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                           hipEvent_t start, hipEvent_t stop) {
    return launch_step_t<false>(dir, half, g, a, stream, start, stop);
}

}  // namespace rtm_hip
