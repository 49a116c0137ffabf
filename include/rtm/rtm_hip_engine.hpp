// rtm_hip_engine.hpp -- HipRTM: the RTMEngine of one MI355X, a thin C++ shell
// over the C ABI in rtm_hip.h (the same adapter a maintainer would add to the
// reference's factory, see src/rtm/factory.cpp:17-38).
//
// Wavefields, snapshots and the image stay on the device across shots; the
// host `snapshots`, `image` and `illumination` arguments of the propagation
// drivers are ignored, as in the reference's CUDA engine
// (src/cuda/propagation_cuda.cu:155-159,232-236), except that a non-null
// `snapshots` pointer in forward_propagation receives a host copy.
#pragma once
#include <string>
#include <vector>
#include "rtm_engine.hpp"

struct rtm_hip_engine;

namespace rtm {

class HipRTM : public RTMEngine {
public:
    // snapshot_format: RTM_HIP_SNAP_*; image_stack: RTM_HIP_STACK_* (0, 0 is
    // the reference's FP32 snapshots and FP32 running stack)
    explicit HipRTM(int device = 0, int snapshot_format = 0, int image_stack = 0);
    ~HipRTM() override;
    HipRTM(const HipRTM&) = delete;
    HipRTM& operator=(const HipRTM&) = delete;

    const char* name() const override;
    bool        is_gpu() const override { return true; }
    int         num_devices() const override { return 1; }
    std::string device_name() const override;

    void setup(const VelocityModel& model, const RTMParams& par, const TimeAxis& ta) override;
    void forward_propagation(const ShotRecord& shot, std::vector<float>* snapshots,
                             std::vector<float>* recorded) override;
    void backward_propagation(const ShotRecord& shot, const std::vector<float>& snapshots,
                              std::vector<float>& image, std::vector<float>& illumination) override;
    // Imaging is fused into the backward step on the device; a host-side call
    // is refused, as in the reference (src/cuda/imaging_cuda.cu:33-37).
    void imaging(const float* fwd_snapshot, const float* bwd_extended, std::vector<float>& image,
                 std::vector<float>& illumination) override;
    void migrate(const std::vector<ShotRecord>& shots, std::vector<float>& image,
                 std::vector<float>& illumination) override;
    void model_shot(ShotRecord& shot) override;

    // Illumination compensation and/or the Laplacian filter
    // (src/rtm/imaging.cpp:48-75) of the image left on the device by the last
    // migrate, computed on the device; flags are RTM_HIP_POST_* (rtm_hip.h).
    // out is assigned nx*nz values, bit-identical to the host functions.
    void postprocess(int flags, float eps_relative, std::vector<float>& out);

    rtm_hip_engine* handle() const { return h_; }

private:
    void check(int status) const;
    void pull_times();
    rtm_hip_engine* h_ = nullptr;
    int snapshot_format_ = 0;
    int image_stack_ = 0;
};

}  // namespace rtm
