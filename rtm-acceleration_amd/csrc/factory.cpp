// factory.cpp -- engine selection by name (the reference's src/rtm/factory.cpp:17-54).
//
//   hip        plain per-shot engine, FP32 snapshots (bit-identical to the
//              reference's CPU image)
//   hip-fp16   FP16 snapshots with one FP32 scale per snapshot, FP64 image stack
//   hip-bf16   BF16 snapshots with one FP32 scale per snapshot, FP64 image stack
//
// The CPU engines of the reference ("cpu", "cpu-opt") are restated under
// oracle/ as the checker and CPU baseline, not as product engines.
#include "rtm/rtm_factory.hpp"

#include <stdexcept>

#include "rtm/rtm_hip_engine.hpp"
#include "rtm_hip.h"

namespace rtm {

std::unique_ptr<RTMEngine> make_engine(const std::string& engine_name, int requested_gpus) {
    if (requested_gpus != 1)
        throw std::runtime_error("this build drives one GPU per engine (--gpus 1)");
    if (engine_name == "hip") return std::make_unique<HipRTM>(0, RTM_HIP_SNAP_F32);
    if (engine_name == "hip-fp16") return std::make_unique<HipRTM>(0, RTM_HIP_SNAP_F16, RTM_HIP_STACK_F64);
    if (engine_name == "hip-bf16") return std::make_unique<HipRTM>(0, RTM_HIP_SNAP_BF16, RTM_HIP_STACK_F64);
    throw std::runtime_error("unknown or unavailable engine: '" + engine_name + "' (see --list-engines)");
}

std::vector<std::string> list_engines() { return {"hip", "hip-fp16", "hip-bf16"}; }

}  // namespace rtm
