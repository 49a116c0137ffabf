// hip_engine.cpp -- HipRTM, the RTMEngine over the C ABI (rtm_hip.h).
#include "rtm/rtm_hip_engine.hpp"

#include <stdexcept>
#include <string>

#include "rtm/benchmark.hpp"
#include "rtm_hip.h"

namespace rtm {

HipRTM::HipRTM(int device, int snapshot_format, int image_stack)
    : snapshot_format_(snapshot_format), image_stack_(image_stack) {
    rtm_hip_engine* h = nullptr;
    const int rc = rtm_hip_create(device, &h);
    if (rc != RTM_HIP_OK)
        throw std::runtime_error(std::string("HIP engine: ") + rtm_hip_last_error(nullptr));
    h_ = h;
}

HipRTM::~HipRTM() { rtm_hip_destroy(h_); }

const char* HipRTM::name() const {
    switch (snapshot_format_) {
        case RTM_HIP_SNAP_F16: return "hip (MI355X fused step, FP16 snapshots)";
        case RTM_HIP_SNAP_BF16: return "hip (MI355X fused step, BF16 snapshots)";
        default: return "hip (MI355X fused step)";
    }
}

std::string HipRTM::device_name() const {
    rtm_hip_info info;
    if (rtm_hip_get_info(h_, &info) != RTM_HIP_OK) return "unknown-gpu";
    return info.device_name;
}

void HipRTM::check(int status) const {
    if (status != RTM_HIP_OK) throw std::runtime_error(rtm_hip_last_error(h_));
}

void HipRTM::pull_times() {
    rtm_hip_times t;
    check(rtm_hip_get_times(h_, &t));
    times.forward = t.forward;
    times.backward = t.backward;
    times.imaging = t.imaging;
    times.h2d = t.h2d;
    times.kernel = t.kernel;
    times.d2h = t.d2h;
}

void HipRTM::setup(const VelocityModel& model, const RTMParams& par, const TimeAxis& ta) {
    par_ = par;
    ta_ = ta;
    g_ = model.grid;
    g_.nb = par.nb;
    g_.half = stencil_half(par.order);
    if (model.v.size() != g_.n_interior())
        throw std::runtime_error("velocity model size does not match nx*nz");
    rtm_hip_grid g{model.grid.nx, model.grid.nz, model.grid.dx, model.grid.dz, model.grid.ox, model.grid.oz};
    rtm_hip_params p{par.nb, par.order, par.sponge_alpha, par.store_interval, par.f0, par.verbose ? 1 : 0,
                     snapshot_format_, image_stack_};
    check(rtm_hip_setup(h_, &g, model.v.data(), &p, ta.nt, ta.dt));
    rtm_hip_info info;
    check(rtm_hip_get_info(h_, &info));
    nsnap_ = info.num_snapshots;
    pull_times();
}

namespace {
rtm_hip_shot view(const ShotRecord& s) {
    return rtm_hip_shot{s.sx, s.sz, s.nrec(), s.rx.data(), s.rz.data(), s.nt, s.dt,
                        s.traces.empty() ? nullptr : s.traces.data()};
}
}  // namespace

void HipRTM::forward_propagation(const ShotRecord& shot, std::vector<float>* snapshots,
                                 std::vector<float>* recorded) {
    const rtm_hip_shot s = view(shot);
    if (snapshots) snapshots->resize(static_cast<std::size_t>(nsnap_) * g_.n_interior());
    if (recorded) recorded->assign(static_cast<std::size_t>(shot.nrec()) * ta_.nt, 0.0f);
    check(rtm_hip_forward(h_, &s, snapshots ? snapshots->data() : nullptr,
                          (recorded && shot.nrec() > 0) ? recorded->data() : nullptr));
    pull_times();
}

void HipRTM::backward_propagation(const ShotRecord& shot, const std::vector<float>& /*snapshots*/,
                                  std::vector<float>& image, std::vector<float>& illumination) {
    const rtm_hip_shot s = view(shot);
    check(rtm_hip_backward(h_, &s));
    // The image lives on the device; mirror it into the caller's vectors so a
    // caller that drives forward/backward itself sees the running sums.
    image.resize(g_.n_interior());
    illumination.resize(g_.n_interior());
    check(rtm_hip_image_get(h_, image.data(), illumination.data()));
    pull_times();
}

void HipRTM::imaging(const float*, const float*, std::vector<float>&, std::vector<float>&) {
    throw std::runtime_error("HipRTM::imaging: the imaging condition is fused into the device backward "
                             "step; call backward_propagation or migrate");
}

void HipRTM::migrate(const std::vector<ShotRecord>& shots, std::vector<float>& image,
                     std::vector<float>& illumination) {
    Timer total;
    std::vector<rtm_hip_shot> v;
    v.reserve(shots.size());
    for (const auto& s : shots) v.push_back(view(s));
    image.assign(g_.n_interior(), 0.0f);
    illumination.assign(g_.n_interior(), 0.0f);
    check(rtm_hip_migrate(h_, static_cast<int>(v.size()), v.data(), image.data(), illumination.data()));
    pull_times();
    times.total += total.elapsed();
}

void HipRTM::model_shot(ShotRecord& shot) {
    shot.nt = ta_.nt;
    shot.dt = ta_.dt;
    std::vector<float> rec(static_cast<std::size_t>(shot.nrec()) * ta_.nt, 0.0f);
    const rtm_hip_shot s = view(shot);
    if (shot.nrec() > 0) check(rtm_hip_model_shot(h_, &s, rec.data()));
    shot.traces.swap(rec);
    pull_times();
}

void HipRTM::postprocess(int flags, float eps_relative, std::vector<float>& out) {
    out.assign(static_cast<std::size_t>(g_.nx) * g_.nz, 0.0f);
    check(rtm_hip_image_postprocess(h_, flags, eps_relative, out.data()));
}

}  // namespace rtm
