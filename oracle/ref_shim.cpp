// ref_shim.cpp -- a C entry layer over the reference engines, compiled
// together with the reference's own sources (taken from where they lie under
// /root/reference by oracle/Makefile) into oracle/_ref/librtmref.so.
//
// TEST INFRASTRUCTURE ONLY: used to pin the oracle restatement bit-for-bit
// against the reference, and timed by bench.py as the CPU baseline.  It calls
// rtm::make_engine / setup / migrate / model_shot (include/rtm_factory.hpp:19-23,
// include/rtm_engine.hpp:20-83) and adds nothing to their arithmetic.
#include <cstdio>
#include <cstring>
#include <exception>
#include <memory>
#include <string>
#include <vector>

#include "rtm_factory.hpp"
#include "rtm_io.hpp"

namespace {
thread_local std::string g_err;

struct ShotIn {            // same layout as orc_shot in rtm_oracle.h
    float sx, sz;
    int nrec;
    const float* rx;
    const float* rz;
    int nt;
    float dt;
    const float* traces;
};
struct GridIn { int nx, nz; float dx, dz, ox, oz; };
struct ParamsIn { int nb, order; float sponge_alpha; int store_interval; float f0; int verbose; };

rtm::ShotRecord to_record(const ShotIn& s) {
    rtm::ShotRecord r;
    r.sx = s.sx; r.sz = s.sz; r.nt = s.nt; r.dt = s.dt;
    r.rx.assign(s.rx, s.rx + s.nrec);
    r.rz.assign(s.rz, s.rz + s.nrec);
    if (s.traces) r.traces.assign(s.traces, s.traces + (std::size_t)s.nrec * s.nt);
    else r.allocate();
    return r;
}

std::unique_ptr<rtm::RTMEngine> make(const char* name, const GridIn* g, const float* v,
                                     const ParamsIn* p, int nt, float dt) {
    auto eng = rtm::make_engine(name);
    rtm::VelocityModel m;
    m.grid.nx = g->nx; m.grid.nz = g->nz; m.grid.dx = g->dx; m.grid.dz = g->dz;
    m.grid.ox = g->ox; m.grid.oz = g->oz;
    m.v.assign(v, v + (std::size_t)g->nx * g->nz);
    rtm::RTMParams par;
    par.nb = p->nb; par.order = p->order; par.sponge_alpha = p->sponge_alpha;
    par.store_interval = p->store_interval; par.f0 = p->f0; par.verbose = p->verbose != 0;
    rtm::TimeAxis ta; ta.nt = nt; ta.dt = dt;
    eng->setup(m, par, ta);
    return eng;
}
}  // namespace

extern "C" {

const char* ref_last_error() { return g_err.c_str(); }

int ref_num_threads(const char* engine) {
    try { return rtm::make_engine(engine)->num_threads(); }
    catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// times: {forward, backward, imaging, total} seconds from engine->times.
int ref_migrate(const char* engine, const GridIn* g, const float* v, const ParamsIn* p,
                int nt, float dt, int nshots, const ShotIn* shots, float* image,
                float* illum, double* times) {
    try {
        auto eng = make(engine, g, v, p, nt, dt);
        std::vector<rtm::ShotRecord> rs;
        for (int i = 0; i < nshots; ++i) rs.push_back(to_record(shots[i]));
        std::vector<float> I, S;
        eng->migrate(rs, I, S);
        std::memcpy(image, I.data(), I.size() * sizeof(float));
        std::memcpy(illum, S.data(), S.size() * sizeof(float));
        if (times) {
            times[0] = eng->times.forward; times[1] = eng->times.backward;
            times[2] = eng->times.imaging; times[3] = eng->times.total;
        }
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// Forward modelling of nshots shots; traces_out holds nshots*nrec*nt floats.
int ref_model_shots(const char* engine, const GridIn* g, const float* v, const ParamsIn* p,
                    int nt, float dt, int nshots, const ShotIn* shots, float* traces_out) {
    try {
        auto eng = make(engine, g, v, p, nt, dt);
        std::size_t off = 0;
        for (int i = 0; i < nshots; ++i) {
            rtm::ShotRecord r = to_record(shots[i]);
            eng->model_shot(r);
            std::memcpy(traces_out + off, r.traces.data(), r.traces.size() * sizeof(float));
            off += r.traces.size();
        }
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

// Forward propagation with host snapshots (nsnap*nx*nz) for snapshot tests.
int ref_forward_snapshots(const char* engine, const GridIn* g, const float* v,
                          const ParamsIn* p, int nt, float dt, const ShotIn* shot,
                          float* snapshots) {
    try {
        auto eng = make(engine, g, v, p, nt, dt);
        std::vector<float> snap((std::size_t)eng->num_snapshots() * g->nx * g->nz, 0.0f);
        rtm::ShotRecord r = to_record(*shot);
        eng->forward_propagation(r, &snap, nullptr);
        std::memcpy(snapshots, snap.data(), snap.size() * sizeof(float));
        return 0;
    } catch (const std::exception& e) { g_err = e.what(); return -1; }
}

void ref_mute_direct_wave(const ShotIn* shot, float* traces, float v, float t_pad,
                          float t_taper) {
    rtm::ShotRecord r = to_record(*shot);
    r.traces.assign(traces, traces + (std::size_t)shot->nrec * shot->nt);
    rtm::mute_direct_wave(r, v, t_pad, t_taper);
    std::memcpy(traces, r.traces.data(), r.traces.size() * sizeof(float));
}

}  // extern "C"
