// step_kernels_s16.hip -- the fused step kernels for 16-bit (FP16/BF16) snapshots
// (step_launch.hpp).
#include "step_launch.hpp"

namespace rtm_hip {

hipError_t launch_step_s16(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                           hipEvent_t start, hipEvent_t stop) {
    return launch_step_t<true>(dir, half, g, a, stream, start, stop);
}

}  // namespace rtm_hip
