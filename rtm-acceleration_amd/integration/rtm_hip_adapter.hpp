// rtm_hip_adapter.hpp -- the MI355X engine as an engine of the reference's
// own C++ tree.
//
// Compiled inside the reference, against its headers (include/rtm_engine.hpp,
// include/rtm_types.hpp) and this repository's C ABI (include/rtm_hip.h); it
// does not build against this repository's C++ API.  HipAdapterRTM is an
// rtm::RTMEngine (include/rtm_engine.hpp:19-84) whose methods forward to the
// rtm_hip_* entry points, so the reference's drivers (src/main.cpp,
// src/tools/make_synthetic.cpp) run the GPU engine unchanged once
// make_engine (src/rtm/factory.cpp:17-38) returns it for "hip".
//
// Engine names accepted by make_hip_adapter: "hip" (FP32 snapshots, the
// reference's image bit for bit), "hip-fp16", "hip-bf16" (16-bit snapshots).
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtm_engine.hpp"
#include "rtm_hip.h"

namespace rtm {

class HipAdapterRTM : public RTMEngine {
public:
    explicit HipAdapterRTM(int snapshot_format = RTM_HIP_SNAP_F32, int device = 0) : format_(snapshot_format) {
        if (rtm_hip_create(device, &h_) != RTM_HIP_OK)
            throw std::runtime_error(std::string("hip engine: ") + rtm_hip_last_error(nullptr));
    }
    ~HipAdapterRTM() override { rtm_hip_destroy(h_); }
    HipAdapterRTM(const HipAdapterRTM&) = delete;
    HipAdapterRTM& operator=(const HipAdapterRTM&) = delete;

    const char* name() const override {
        return format_ == RTM_HIP_SNAP_F16    ? "hip (MI355X, FP16 snapshots)"
               : format_ == RTM_HIP_SNAP_BF16 ? "hip (MI355X, BF16 snapshots)"
                                              : "hip (MI355X)";
    }
    bool is_gpu() const override { return true; }
    int num_devices() const override { return 1; }
    std::string device_name() const override {
        rtm_hip_info info;
        return rtm_hip_get_info(h_, &info) == RTM_HIP_OK ? std::string(info.device_name) : "unknown-gpu";
    }

    void setup(const VelocityModel& m, const RTMParams& p, const TimeAxis& ta) override {
        par_ = p;
        ta_ = ta;
        g_ = m.grid;
        g_.nb = p.nb;
        g_.half = stencil_half(p.order);
        const rtm_hip_grid g{m.grid.nx, m.grid.nz, m.grid.dx, m.grid.dz, m.grid.ox, m.grid.oz};
        const rtm_hip_params hp{p.nb, p.order, p.sponge_alpha, p.store_interval, p.f0, p.verbose ? 1 : 0,
                                format_, RTM_HIP_STACK_F32};
        check(rtm_hip_setup(h_, &g, m.v.data(), &hp, ta.nt, ta.dt));
        rtm_hip_info info;
        check(rtm_hip_get_info(h_, &info));
        nsnap_ = info.num_snapshots;
        pull_times();
    }

    // snapshots stay on the device unless the caller asks for them
    void forward_propagation(const ShotRecord& s, std::vector<float>* snaps, std::vector<float>* rec) override {
        const rtm_hip_shot v = view(s);
        if (snaps) snaps->assign((size_t)nsnap_ * g_.n_interior(), 0.0f);
        if (rec) rec->assign((size_t)s.nrec() * ta_.nt, 0.0f);
        check(rtm_hip_forward(h_, &v, snaps ? snaps->data() : nullptr, rec ? rec->data() : nullptr));
        pull_times();
    }

    // accumulates into the device image of the preceding forward; the host
    // image and illumination are filled by migrate
    void backward_propagation(const ShotRecord& s, const std::vector<float>&, std::vector<float>& image,
                              std::vector<float>& illum) override {
        const rtm_hip_shot v = view(s);
        check(rtm_hip_backward(h_, &v));
        image.resize(g_.n_interior());
        illum.resize(g_.n_interior());
        check(rtm_hip_image_get(h_, image.data(), illum.data()));
        pull_times();
    }

    void imaging(const float*, const float*, std::vector<float>&, std::vector<float>&) override {
        throw std::runtime_error("hip engine: imaging is fused into the device backward step");
    }

    void migrate(const std::vector<ShotRecord>& shots, std::vector<float>& image,
                 std::vector<float>& illum) override {
        std::vector<rtm_hip_shot> v;
        v.reserve(shots.size());
        for (const ShotRecord& s : shots) v.push_back(view(s));
        image.assign(g_.n_interior(), 0.0f);
        illum.assign(g_.n_interior(), 0.0f);
        check(rtm_hip_migrate(h_, (int)v.size(), v.data(), image.data(), illum.data()));
        pull_times();
    }

    void model_shot(ShotRecord& shot) override {
        shot.nt = ta_.nt;
        shot.dt = ta_.dt;
        std::vector<float> rec((size_t)shot.nrec() * ta_.nt, 0.0f);
        shot.traces.clear();
        const rtm_hip_shot v = view(shot);
        check(rtm_hip_model_shot(h_, &v, rec.data()));
        shot.traces = std::move(rec);
        pull_times();
    }

private:
    static rtm_hip_shot view(const ShotRecord& s) {
        return rtm_hip_shot{s.sx, s.sz, s.nrec(), s.rx.data(), s.rz.data(), s.nt, s.dt,
                            s.traces.empty() ? nullptr : s.traces.data()};
    }
    void check(int rc) const {
        if (rc != RTM_HIP_OK) throw std::runtime_error(std::string("hip engine: ") + rtm_hip_last_error(h_));
    }
    // the driver owns times.io and times.total; the engine fills the rest
    void pull_times() {
        rtm_hip_times t;
        check(rtm_hip_get_times(h_, &t));
        times.forward = t.forward;
        times.backward = t.backward;
        times.imaging = t.imaging;
        times.h2d = t.h2d;
        times.kernel = t.kernel;
        times.d2h = t.d2h;
    }

    rtm_hip_engine* h_ = nullptr;
    int format_;
};

// The factory hook: the engine for a "hip*" name, or null for any other.
inline std::unique_ptr<RTMEngine> make_hip_adapter(const std::string& name) {
    if (name == "hip") return std::make_unique<HipAdapterRTM>(RTM_HIP_SNAP_F32);
    if (name == "hip-fp16") return std::make_unique<HipAdapterRTM>(RTM_HIP_SNAP_F16);
    if (name == "hip-bf16") return std::make_unique<HipAdapterRTM>(RTM_HIP_SNAP_BF16);
    return nullptr;
}

}  // namespace rtm
