// device_engine.hip -- the C ABI of include/rtm_hip.h: one handle = one shot
// engine on one device, everything resident in HBM.
//
// Device layout (per handle):
//   vdt2, three wavefields      [nxe][ldz] float, ldz = nze rounded up to 64
//                               (256-byte rows; pad cells stay 0)
//   wx [nxe], wz [ldz]          separable Cerjan tapers
//   wavelet [nt]                Ricker source
//   snapshot store              num_snapshots x [nx][nz] float (the reference's
//                               interior layout, src/rtm/forward.cpp:83-88)
//   image, illumination         [nx][nz] float, running sums over shots
//   traces                      [nrec][nt_stride] float, uploaded per shot
//   per-tile entry tables       injection / recording cells, built per shot
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "host_tables.hpp"
#include "rtm/benchmark.hpp"
#include "rtm/rtm_types.hpp"
#include "rtm_hip.h"
#include "image_post.hpp"
#include "step_kernels.hpp"

using rtm_hip::Dir;
using rtm_hip::Entry;

// Scratch of rtm_hip_image_postprocess_noalloc: the partial maxima of S, the
// only device memory the post-processing kernels use beyond their inputs and
// output.  kNoAllocScratchFloats floats (2 KiB) per GPU, held until the
// process ends; every call on a GPU shares that GPU's copy (calls that
// overlap on one GPU race) and writes every partial it reads, so nothing is
// carried over from the previous call.  Host code uses its device address
// (rtm_hip_engine::d_scratch, resolved once per handle in rtm_hip_create),
// never this name.
constexpr int kNoAllocScratchFloats = rtm_hip::kPostMaxParts;
static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

namespace {

thread_local std::string g_null_handle_error;

struct EventPair {
    hipEvent_t start = nullptr, stop = nullptr;
};

int round_up(int v, int m) { return (v + m - 1) / m * m; }

}  // namespace

struct rtm_hip_engine {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    bool ready = false;

    // geometry / parameters
    int nx = 0, nz = 0, nb = 0, half = 0, order = 0, nxe = 0, nze = 0, ldz = 0;
    int nt = 0, store = 1, nsnap = 0, snap_fmt = 0, verbose = 0;
    float dx = 0, dz = 0, ox = 0, oz = 0, dt = 0;
    size_t n_int = 0, n_fld = 0;
    rtm_hip::Coeffs coef{};
    int family_user = -1;     // -1: built-in choice
    int tile_x_user = 0;      // 0: built-in choice
    bool tuned[2] = {false, false};            // per direction: forward, backward
    rtm_hip::StepGeometry tuned_geom[2];

    // device buffers
    float* d_vdt2 = nullptr;
    float* d_wx = nullptr;
    float* d_wz = nullptr;
    float* d_wavelet = nullptr;
    float* d_fld[2] = {nullptr, nullptr};     // p^{n-1} (overwritten by p^{n+1}) and p^n
    float* d_hold[2] = {nullptr, nullptr};   // fields held for paired imaging (pair_step)
    float* d_snap = nullptr;         // FP32 snapshots
    uint16_t* d_snap16 = nullptr;    // FP16/BF16 snapshot codes
    float* d_snap_max = nullptr;     // per-snapshot max|p| (16-bit formats)
    // scale bookkeeping of 16-bit snapshots (StepArgs), double-buffered by
    // snapshot parity: per-wave maxima [2][wave_max_cap], their slice maxima
    // [2][kReduceWaves], and the counts {waves[2], slices[2]}
    float* d_wave_max = nullptr;
    int wave_max_cap = 0;
    float* d_slice_max = nullptr;
    int* d_part_counts = nullptr;
    size_t snap_bytes = 0;
    float* d_image = nullptr;
    float* d_illum = nullptr;
    int image_stack = 0;             // 0: FP32 running sum, 1: FP64 stack
    double* d_stack_i = nullptr;
    double* d_stack_s = nullptr;
    float* d_post = nullptr;         // post-processing scratch + output (allocated on first use)
    float* d_scratch = nullptr;      // this device's copy of `scratch` (rtm_hip_image_postprocess_noalloc)
    size_t post_cap = 0;
    int* d_tile_ptr[2] = {nullptr, nullptr};
    size_t tile_ptr_cap[2] = {0, 0};
    Entry* d_ent[2] = {nullptr, nullptr};
    size_t ent_cap[2] = {0, 0};
    float* d_traces = nullptr;
    size_t traces_cap = 0;
    float* d_amps = nullptr;         // receiver amplitudes in backward-entry order (StepArgs::amps)
    size_t amps_cap = 0;
    float* d_rec = nullptr;
    size_t rec_cap = 0;
    size_t device_bytes = 0;

    // host staging kept alive until the stream has consumed it
    std::vector<int> h_tile_ptr[2];
    std::vector<Entry> h_ent[2];

    // accounting
    rtm_hip_times times{};
    int timing_every = 0;
    long long launches[2] = {0, 0};
    long long sampled[2] = {0, 0};
    double sampled_ms[2] = {0, 0};
    std::vector<EventPair> ev_pool;
    struct PendingEvents {
        int kind;            // 0 forward, 1 backward
        EventPair ev;
        long long launches;  // step launches between the two events
    };
    std::vector<PendingEvents> ev_pending;
    EventPair loop_ev;

    ~rtm_hip_engine() { release(); }

    int fail(int code, const char* fmt, ...) {
        char buf[1024];
        va_list ap;
        va_start(ap, fmt);
        std::vsnprintf(buf, sizeof buf, fmt, ap);
        va_end(ap);
        err = buf;
        return code;
    }

    int hip_fail(hipError_t e, const char* what) {
        return fail(e == hipErrorOutOfMemory ? RTM_HIP_ERR_MEMORY : RTM_HIP_ERR_DEVICE,
                    "%s: %s (%s)", what, hipGetErrorString(e), hipGetErrorName(e));
    }

    template <typename T>
    int alloc(T** p, size_t count, const char* what) {
        if (count == 0) count = 1;
        hipError_t e = hipMalloc(reinterpret_cast<void**>(p), count * sizeof(T));
        if (e != hipSuccess) {
            *p = nullptr;
            return hip_fail(e, what);
        }
        device_bytes += count * sizeof(T);
        return RTM_HIP_OK;
    }

    template <typename T>
    int grow(T** p, size_t* cap, size_t count, const char* what) {
        if (count <= *cap && *p) return RTM_HIP_OK;
        if (*p) {
            (void)hipFree(*p);
            device_bytes -= *cap * sizeof(T);
            *p = nullptr;
            *cap = 0;
        }
        const size_t n = std::max<size_t>(count, 1);
        if (int rc = alloc(p, n, what)) return rc;
        *cap = n;
        return RTM_HIP_OK;
    }

    void release() {
        if (device >= 0) (void)hipSetDevice(device);
        auto f = [](auto*& p) {
            if (p) (void)hipFree(p);
            p = nullptr;
        };
        f(d_vdt2); f(d_wx); f(d_wz); f(d_wavelet);
        for (auto& p : d_fld) f(p);
        for (auto& p : d_hold) f(p);
        f(d_snap); f(d_snap16); f(d_snap_max); f(d_wave_max); f(d_slice_max); f(d_part_counts); f(d_image); f(d_illum); f(d_traces); f(d_amps); f(d_rec);
        f(d_stack_i); f(d_stack_s); f(d_post);
        post_cap = 0;
        for (int k = 0; k < 2; ++k) {
            f(d_tile_ptr[k]); f(d_ent[k]);
            tile_ptr_cap[k] = ent_cap[k] = 0;
        }
        traces_cap = amps_cap = rec_cap = 0;
        device_bytes = 0;
        for (auto& ev : ev_pool) {
            (void)hipEventDestroy(ev.start);
            (void)hipEventDestroy(ev.stop);
        }
        ev_pool.clear();
        for (auto& pe : ev_pending) {
            (void)hipEventDestroy(pe.ev.start);
            (void)hipEventDestroy(pe.ev.stop);
        }
        ev_pending.clear();
        if (loop_ev.start) (void)hipEventDestroy(loop_ev.start);
        if (loop_ev.stop) (void)hipEventDestroy(loop_ev.stop);
        loop_ev = EventPair{};
        ready = false;
    }

    // Step-kernel tiling of direction k (0 forward, 1 backward): the caller's
    // choice, else the autotuned one, else the DPP march mapping with 8-column
    // tiles (4 on small grids, so the launch still spreads over every CU).
    rtm_hip::StepGeometry geometry(int k) const {
        if (family_user < 0 && tile_x_user == 0 && tuned[k]) return tuned_geom[k];
        const auto fam = family_user < 0 ? rtm_hip::Family::kMarchDpp : static_cast<rtm_hip::Family>(family_user);
        int tx = tile_x_user;
        if (tx == 0 && fam == rtm_hip::Family::kStream) tx = 16;
        if (tx == 0) tx = (long long)nxe * ldz >= 4LL * 1024 * 1024 ? 8 : 4;
        return rtm_hip::step_geometry(fam, half, tx, nxe, nze, ldz);
    }
};

namespace {

// Nearest grid point with clamping to the interior (src/rtm/rtm_cpu.cpp:99-123).
struct CellIndex {
    int ixe, ize;
};

CellIndex to_cell(const rtm_hip_engine* h, float x, float z, int* nclamped) {
    const int ix0 = static_cast<int>(std::lround((x - h->ox) / h->dx));
    const int iz0 = static_cast<int>(std::lround((z - h->oz) / h->dz));
    const int ix = std::clamp(ix0, 0, h->nx - 1);
    const int iz = std::clamp(iz0, 0, h->nz - 1);
    if (ix != ix0 || iz != iz0) ++*nclamped;
    return {ix + h->nb, iz + h->nb};
}

// Per-tile entry table for one direction.  Entries are ordered by tile, then
// cell, then code, so at a cell the source (code -1) precedes recordings and
// duplicate receivers keep ascending receiver order (src/rtm/backward.cpp:23-26).
int build_entries(rtm_hip_engine* h, int k, const rtm_hip_shot* s, bool with_source,
                  bool with_receivers, const rtm_hip::StepGeometry& g) {
    const int ntiles = g.ntiles;
    int nclamped = 0;
    struct Tagged {
        int tile, cell, code;
    };
    std::vector<Tagged> v;
    v.reserve(static_cast<size_t>(s->nrec) + 1);
    auto add = [&](CellIndex c, int code) {
        const int tile = rtm_hip::tile_of_cell(g, c.ixe, c.ize);
        v.push_back({tile, c.ixe * h->ldz + c.ize, code});
    };
    const CellIndex src = to_cell(h, s->sx, s->sz, &nclamped);
    if (with_source) add(src, -1);
    if (with_receivers)
        for (int ir = 0; ir < s->nrec; ++ir) add(to_cell(h, s->rx[ir], s->rz[ir], &nclamped), ir);
    else
        for (int ir = 0; ir < s->nrec; ++ir) (void)to_cell(h, s->rx[ir], s->rz[ir], &nclamped);
    std::sort(v.begin(), v.end(), [](const Tagged& a, const Tagged& b) {
        if (a.tile != b.tile) return a.tile < b.tile;
        if (a.cell != b.cell) return a.cell < b.cell;
        return a.code < b.code;
    });
    if (nclamped > 0 && h->verbose && k == 0)
        std::fprintf(stderr,
                     "WARNING: %d of %d source/receiver positions fell outside the "
                     "%.1f x %.1f m model and were clamped to the edge.\n",
                     nclamped, s->nrec + 1, (h->nx - 1) * h->dx, (h->nz - 1) * h->dz);

    auto& ptr = h->h_tile_ptr[k];
    auto& ent = h->h_ent[k];
    ptr.assign(static_cast<size_t>(ntiles) + 1, 0);
    ent.resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        ent[i] = Entry{v[i].cell, v[i].code};
        ++ptr[static_cast<size_t>(v[i].tile) + 1];
    }
    for (int t = 0; t < ntiles; ++t) ptr[t + 1] += ptr[t];

    if (int rc = h->grow(&h->d_tile_ptr[k], &h->tile_ptr_cap[k], ptr.size(), "tile table"))
        return rc;
    if (int rc = h->grow(&h->d_ent[k], &h->ent_cap[k], std::max<size_t>(ent.size(), 1), "entry table"))
        return rc;
    rtm::Timer t;
    hipError_t e = hipMemcpyAsync(h->d_tile_ptr[k], ptr.data(), ptr.size() * sizeof(int),
                                  hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess && !ent.empty())
        e = hipMemcpyAsync(h->d_ent[k], ent.data(), ent.size() * sizeof(Entry),
                           hipMemcpyHostToDevice, h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "upload of injection tables");
    h->times.h2d += t.elapsed();
    return RTM_HIP_OK;
}

EventPair take_event(rtm_hip_engine* h) {
    if (!h->ev_pool.empty()) {
        EventPair p = h->ev_pool.back();
        h->ev_pool.pop_back();
        return p;
    }
    EventPair p;
    (void)hipEventCreate(&p.start);
    (void)hipEventCreate(&p.stop);
    return p;
}

// Reads back the sampled kernel events once the stream is idle.
void collect_events(rtm_hip_engine* h) {
    for (auto& pe : h->ev_pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pe.ev.start, pe.ev.stop) == hipSuccess) {
            h->sampled_ms[pe.kind] += ms;
            h->sampled[pe.kind] += pe.launches;
        }
        h->ev_pool.push_back(pe.ev);
    }
    h->ev_pending.clear();
}

int check_handle(const rtm_hip_engine* h, bool need_setup) {
    if (!h) {
        g_null_handle_error = "null rtm_hip_engine handle";
        return RTM_HIP_ERR_ARG;
    }
    if (need_setup && !h->ready) {
        const_cast<rtm_hip_engine*>(h)->err = "engine is not set up (call rtm_hip_setup first)";
        return RTM_HIP_ERR_ARG;
    }
    return RTM_HIP_OK;
}

int validate_shot(rtm_hip_engine* h, const rtm_hip_shot* s, bool need_traces) {
    if (!s) return h->fail(RTM_HIP_ERR_ARG, "null shot");
    if (s->nrec < 0) return h->fail(RTM_HIP_ERR_ARG, "negative receiver count");
    if (s->nrec > 0 && (!s->rx || !s->rz)) return h->fail(RTM_HIP_ERR_ARG, "null receiver coordinates");
    if (need_traces) {
        if (s->nrec > 0 && !s->traces) return h->fail(RTM_HIP_ERR_ARG, "backward propagation needs traces");
        if (s->nt < h->nt)
            return h->fail(RTM_HIP_ERR_ARG, "shot has %d samples per trace, fewer than the engine's nt=%d",
                           s->nt, h->nt);
    }
    return RTM_HIP_OK;
}

// Snapshot operands of the launch for time index it (see StepArgs).
void set_snapshot_args(const rtm_hip_engine* h, Dir dir, int it, rtm_hip::StepArgs& a) {
    a.snap_fmt = h->snap_fmt;
    a.snap_out = nullptr;
    a.wave_max = nullptr;
    a.wave_count = nullptr;
    a.red_in = nullptr;
    a.red_count = nullptr;
    a.red_out = nullptr;
    a.red_out_count = nullptr;
    a.quant_parts = nullptr;
    a.quant_count = nullptr;
    a.quant_out = nullptr;
    a.quant_max_out = nullptr;
    a.parts_cap = h->wave_max_cap;
    a.snap_in = nullptr;
    a.snap16_in = nullptr;
    a.snap16_max = nullptr;
    a.pending_field = nullptr;
    a.hold_field = nullptr;
    a.snap_in2 = nullptr;
    a.snap16_in2 = nullptr;
    a.snap16_max2 = nullptr;
    const int st = h->store;
    const bool stored = (it % st) == 0;
    const size_t slot = (size_t)(it / st);
    if (dir == Dir::Backward && stored) {
        // paired imaging: a Hold step images nothing (its field goes to a
        // hold buffer, run_loop), a Pair step images the held stored step
        // (slot + 1) first, then its own
        const rtm_hip::PairStep p = rtm_hip::pair_step(it, st, rtm_hip::pair_first(h->nt, st, false));
        if (p.role == rtm_hip::PairRole::Hold) return;
        if (p.role == rtm_hip::PairRole::Pair) {
            a.hold_field = h->d_hold[p.hold];
            if (h->snap_fmt == RTM_HIP_SNAP_F32) {
                a.snap_in2 = h->d_snap + (slot + 1) * h->n_int;
            } else {
                a.snap16_in2 = h->d_snap16 + (slot + 1) * h->n_int;
                a.snap16_max2 = h->d_snap_max + slot + 1;
            }
        }
    }
    if (h->snap_fmt == RTM_HIP_SNAP_F32) {
        if (!stored) return;
        if (dir == Dir::Forward) a.snap_out = h->d_snap + slot * h->n_int;
        else a.snap_in = h->d_snap + slot * h->n_int;
        return;
    }
    constexpr int kRW = rtm_hip::kReduceWaves;
    if (dir == Dir::Forward) {
        if (stored) {                                   // per-wave maxima of the new field
            const int p = (int)(slot & 1);
            a.wave_max = h->d_wave_max + (size_t)p * h->wave_max_cap;
            a.wave_count = h->d_part_counts + p;
        }
        if (it >= 1 && ((it - 1) % st) == 0) {          // their slices, one step later
            const int p = (int)(((it - 1) / st) & 1);
            a.red_in = h->d_wave_max + (size_t)p * h->wave_max_cap;
            a.red_count = h->d_part_counts + p;
            a.red_out = h->d_slice_max + p * kRW;
            a.red_out_count = h->d_part_counts + 2 + p;
        }
        if (it >= 2 && ((it - 2) % st) == 0) {          // the codes, two steps later
            const size_t s2 = (size_t)((it - 2) / st);
            const int p = (int)(s2 & 1);
            a.quant_parts = h->d_slice_max + p * kRW;
            a.quant_count = h->d_part_counts + 2 + p;
            a.quant_out = h->d_snap16 + s2 * h->n_int;
            a.quant_max_out = h->d_snap_max + s2;
        }
    } else if (stored) {
        a.snap16_in = h->d_snap16 + slot * h->n_int;
        a.snap16_max = h->d_snap_max + slot;
    }
}

// Runs one propagation (forward or backward) of h->nt steps.
// snapshots false: a forward pass that stores no snapshots (modelling, as
// the reference's model_shot passes no snapshot store, src/rtm/rtm.cpp:109):
// the step arguments keep their null snapshot operands and FP32 format.
int run_loop(rtm_hip_engine* h, Dir dir, const rtm_hip_shot* s, bool record,
             const rtm_hip::StepGeometry& g, bool snapshots = true) {
    const int k = dir == Dir::Forward ? 0 : 1;
    rtm_hip::StepArgs a{};
    a.vdt2 = h->d_vdt2;
    a.wx = h->d_wx;
    a.wz = h->d_wz;
    a.coef = h->coef;
    a.nxe = h->nxe; a.nze = h->nze; a.ldz = h->ldz; a.nb = h->nb; a.nx = h->nx; a.nz = h->nz;
    a.ntz = g.ntz;
    a.ntz_magic = g.ntz_magic;
    a.ntx = g.ntx;
    a.ntiles = g.ntiles;
    a.tile_ptr = h->d_tile_ptr[k];
    a.entries = h->d_ent[k];
    a.wavelet = h->d_wavelet;
    a.amps = h->d_amps;
    a.amp_count = (int)h->h_ent[1].size();
    a.rec_out = record ? h->d_rec : nullptr;
    a.rec_stride = h->nt;
    a.image = h->d_image;
    a.illum = h->d_illum;

    for (auto& p : h->d_fld) {
        hipError_t e = hipMemsetAsync(p, 0, h->n_fld * sizeof(float), h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "wavefield reset");
    }
    if (dir == Dir::Forward && snapshots && h->d_snap_max) {
        hipError_t e = hipMemsetAsync(h->d_snap_max, 0, (size_t)h->nsnap * sizeof(float), h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "snapshot scale reset");
    }
    if (!h->loop_ev.start) {
        (void)hipEventCreate(&h->loop_ev.start);
        (void)hipEventCreate(&h->loop_ev.stop);
    }
    (void)hipEventRecord(h->loop_ev.start, h->stream);

    // every < 0: one event pair around all step launches of the loop
    EventPair phase_ev;
    if (h->timing_every < 0) {
        phase_ev = take_event(h);
        (void)hipEventRecord(phase_ev.start, h->stream);
    }
    // Two wavefield buffers: each step writes p^{n+1} over p^{n-1}, in place.
    // A step reads p^{n-1} only at its own output cells, in the lane that
    // writes p^{n+1} there, and before that write (the value feeds the
    // store), so no other wave needs the old value; the step's working set is
    // three fields instead of four, which on a grid of a few million cells
    // stays within the L2 caches from one launch to the next.  On a Hold step
    // of paired imaging the field goes to a hold buffer instead, which the
    // following steps read as p^n and p^{n-1} and the Pair step as the held
    // field; the step after it, whose p^{n-1} is held, writes the rotation
    // buffer that does not hold its p^n.  Both buffers start zeroed, so steps
    // 0 and 1 read zeros.
    float* const r0 = h->d_fld[0];
    float* const r1 = h->d_fld[1];
    float* prev = r0;
    float* cur = r1;
    const int first = rtm_hip::pair_first(h->nt, h->store, false);
    for (int step = 0; step < h->nt; ++step) {
        const int it = dir == Dir::Forward ? step : h->nt - 1 - step;
        a.it = it;
        float* out = prev;
        if (dir == Dir::Backward && snapshots) {
            const rtm_hip::PairStep p = rtm_hip::pair_step(it, h->store, first);
            if (p.role == rtm_hip::PairRole::Hold) out = h->d_hold[p.hold];
            else if (prev != r0 && prev != r1) out = cur == r0 ? r1 : r0;
        }
        a.p_prev = prev;
        a.p_cur = cur;
        a.p_next = out;
        if (snapshots) set_snapshot_args(h, dir, it, a);

        const bool sample = h->timing_every > 0 && (h->launches[k] % h->timing_every) == 0;
        EventPair ev;
        if (sample) ev = take_event(h);
        hipError_t e = rtm_hip::launch_step(dir, h->half, g, a, h->stream, ev.start, ev.stop);
        if (e != hipSuccess) return h->hip_fail(e, "step kernel launch");
        if (sample) h->ev_pending.push_back({k, ev, 1});
        h->launches[k] += 1;
        prev = cur;
        cur = out;
    }
    if (phase_ev.start) {
        (void)hipEventRecord(phase_ev.stop, h->stream);
        h->ev_pending.push_back({k, phase_ev, h->nt});
    }
    // 16-bit stored steps among the last two have no step it+2 in the loop:
    // their max from the slices (step nt-2) or the per-wave maxima (step
    // nt-1), then their codes from the field (after the loop, cur / prev hold
    // the fields of steps nt-1 / nt-2)
    if (dir == Dir::Forward && snapshots && h->snap_fmt != RTM_HIP_SNAP_F32) {
        const int nparts = g.family == rtm_hip::Family::kLdsTile ? g.ntiles * rtm_hip::kWaves : g.ntiles;
        for (int it = std::max(0, h->nt - 2); it < h->nt; ++it) {
            if (it % h->store) continue;
            const size_t slot = (size_t)(it / h->store);
            const int p = (int)(slot & 1);
            const bool last = it == h->nt - 1;
            hipError_t e = last ? rtm_hip::launch_reduce_max(h->d_wave_max + (size_t)p * h->wave_max_cap, nparts,
                                                             h->d_snap_max + slot, h->stream)
                                : rtm_hip::launch_reduce_max(h->d_slice_max + p * rtm_hip::kReduceWaves,
                                                             std::min(rtm_hip::kReduceWaves, nparts),
                                                             h->d_snap_max + slot, h->stream);
            if (e == hipSuccess)
                e = rtm_hip::launch_quantise(last ? cur : prev, h->ldz, h->nb, h->nx, h->nz,
                                             static_cast<rtm_hip::SnapFmt>(h->snap_fmt), h->d_snap_max + slot,
                                             h->d_snap16 + slot * h->n_int, h->stream);
            if (e != hipSuccess) return h->hip_fail(e, "snapshot quantisation");
        }
    }
    if (dir == Dir::Backward && h->image_stack) {
        hipError_t e = rtm_hip::launch_fold_fp64(h->d_image, h->d_illum, h->d_stack_i, h->d_stack_s, h->n_int,
                                                 h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "FP64 stack update");
    }
    (void)hipEventRecord(h->loop_ev.stop, h->stream);
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return h->hip_fail(e, "step loop");
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, h->loop_ev.start, h->loop_ev.stop) == hipSuccess)
        h->times.kernel += 1e-3 * ms;
    collect_events(h);
    return RTM_HIP_OK;
}

// Autotuner: times every step-kernel family and tile width on this grid for
// one direction (best of 3 x 48 launches each, no injection entries, imaging or snapshot
// traffic on every store-th launch as in the real loop) and keeps the
// fastest.  Runs once per setup and direction, before the direction's first
// propagation; the wavefields are reset afterwards by the propagation itself.
int autotune(rtm_hip_engine* h, Dir dir) {
    const int k = dir == Dir::Forward ? 0 : 1;
    if (h->tuned[k] || h->family_user >= 0 || h->tile_x_user != 0) return RTM_HIP_OK;
    struct Candidate {
        rtm_hip::Family family;
        int tile_x;
    };
    std::vector<Candidate> cands;
    for (auto f : {rtm_hip::Family::kMarchDpp, rtm_hip::Family::kStream, rtm_hip::Family::kMarch,
                   rtm_hip::Family::kLdsTile})
        for (int t : rtm_hip::tile_widths(f)) cands.push_back({f, t});
    int max_tiles = 0;
    for (const Candidate& c : cands)
        max_tiles = std::max(max_tiles,
                             rtm_hip::step_geometry(c.family, h->half, c.tile_x, h->nxe, h->nze, h->ldz).ntiles);
    int* d_ptr = nullptr;
    float* d_img = nullptr;
    if (hipMalloc(&d_ptr, (size_t)(max_tiles + 1) * sizeof(int)) != hipSuccess ||
        hipMalloc(&d_img, 2 * h->n_int * sizeof(float)) != hipSuccess ||
        hipMemsetAsync(d_ptr, 0, (size_t)(max_tiles + 1) * sizeof(int), h->stream) != hipSuccess ||
        hipMemsetAsync(d_img, 0, 2 * h->n_int * sizeof(float), h->stream) != hipSuccess) {
        if (d_ptr) (void)hipFree(d_ptr);
        if (d_img) (void)hipFree(d_img);
        return h->fail(RTM_HIP_ERR_MEMORY, "autotune scratch allocation failed");
    }
    EventPair ev = take_event(h);
    rtm_hip::StepArgs a{};
    a.vdt2 = h->d_vdt2; a.wx = h->d_wx; a.wz = h->d_wz; a.coef = h->coef;
    a.nxe = h->nxe; a.nze = h->nze; a.ldz = h->ldz; a.nb = h->nb; a.nx = h->nx; a.nz = h->nz;
    a.tile_ptr = d_ptr; a.entries = nullptr; a.wavelet = h->d_wavelet;
    a.image = d_img; a.illum = d_img + h->n_int;
    float best_ms = 1e30f;
    rtm_hip::StepGeometry best = h->geometry(k);
    int rc = RTM_HIP_OK;
    // per candidate: 4 warm launches, then the best of 3 rounds of 48 timed
    // launches (the minimum is the least noisy estimate of a kernel's time)
    constexpr int kWarm = 4, kRounds = 3, kTimed = 48;
    auto time_geometry = [&](const rtm_hip::StepGeometry& g) -> float {
        a.ntz = g.ntz; a.ntz_magic = g.ntz_magic; a.ntx = g.ntx; a.ntiles = g.ntiles;
        float cand_ms = 1e30f;
        int n = 0;
        for (int round = 0; round < kRounds && !rc; ++round) {
            const int count = (round == 0 ? kWarm : 0) + kTimed;
            for (int i = 0; i < count && !rc; ++i, ++n) {
                if (i == count - kTimed) (void)hipEventRecord(ev.start, h->stream);
                // the loop's in-place update over two buffers
                a.p_prev = h->d_fld[n & 1];
                a.p_cur = h->d_fld[(n + 1) & 1];
                a.p_next = h->d_fld[n & 1];
                // slot 0 (and its neighbour for 16-bit quantisation) stand in
                // for the real slots; the propagation rewrites them
                set_snapshot_args(h, dir, n % std::min(h->nt, 2 * h->store), a);
                hipError_t e = rtm_hip::launch_step(dir, h->half, g, a, h->stream);
                if (e != hipSuccess) rc = h->hip_fail(e, "autotune launch");
            }
            (void)hipEventRecord(ev.stop, h->stream);
            if (rc) break;
            if (hipEventSynchronize(ev.stop) != hipSuccess) {
                rc = h->fail(RTM_HIP_ERR_DEVICE, "autotune synchronisation failed");
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, ev.start, ev.stop);
            cand_ms = std::min(cand_ms, ms);
        }
        return cand_ms;
    };
    // stage 1: mapping and tile width at full occupancy; candidates are
    // tried in order of preference and a later one must win by 2 % so that
    // timing noise does not flip the choice (48 launches per round keep the
    // rounds' spread below that)
    for (const Candidate& c : cands) {
        const rtm_hip::StepGeometry g = rtm_hip::step_geometry(c.family, h->half, c.tile_x, h->nxe, h->nze, h->ldz);
        const float ms = time_geometry(g);
        if (rc) break;
        if (ms < 0.98f * best_ms) {
            best_ms = ms;
            best = g;
        }
    }
    // stage 2: wavefronts per CU for the winner (fewer resident workgroups
    // leave more L2 to each)
    if (!rc && best.family != rtm_hip::Family::kLdsTile) {
        const rtm_hip::StepGeometry base = best;
        for (int cap : {4, 2}) {
            rtm_hip::StepGeometry g = base;
            g.blocks_per_cu = cap;
            g.lds_bytes = rtm_hip::occupancy_lds_bytes(h->device, cap);
            if (g.lds_bytes == 0) continue;
            const float ms = time_geometry(g);
            if (rc) break;
            if (ms < 0.95f * best_ms) {
                best_ms = ms;
                best = g;
            }
        }
    }
    h->ev_pool.push_back(ev);
    (void)hipStreamSynchronize(h->stream);
    (void)hipFree(d_ptr);
    (void)hipFree(d_img);
    if (rc) return rc;
    h->tuned[k] = true;
    h->tuned_geom[k] = best;
    return RTM_HIP_OK;
}

int forward_impl(rtm_hip_engine* h, const rtm_hip_shot* s, float* snapshots_out, float* recorded,
                 bool snapshots = true) {
    if (int rc = validate_shot(h, s, false)) return rc;
    rtm::Timer wall;
    if (int rc = autotune(h, Dir::Forward)) return rc;
    const rtm_hip::StepGeometry g = h->geometry(0);
    const bool record = recorded != nullptr && s->nrec > 0;
    if (int rc = build_entries(h, 0, s, true, record, g)) return rc;
    if (record) {
        const size_t n = (size_t)s->nrec * h->nt;
        if (int rc = h->grow(&h->d_rec, &h->rec_cap, n, "recorded traces")) return rc;
        hipError_t e = hipMemsetAsync(h->d_rec, 0, n * sizeof(float), h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "recorded traces reset");
    }
    if (int rc = run_loop(h, Dir::Forward, s, record, g, snapshots)) return rc;
    h->times.forward += wall.elapsed();

    rtm::Timer d2h;
    if (record) {
        hipError_t e = hipMemcpy(recorded, h->d_rec, (size_t)s->nrec * h->nt * sizeof(float),
                                 hipMemcpyDeviceToHost);
        if (e != hipSuccess) return h->hip_fail(e, "download of recorded traces");
    }
    if (snapshots_out) {
        const size_t n = (size_t)h->nsnap * h->n_int;
        if (h->snap_fmt == RTM_HIP_SNAP_F32) {
            hipError_t e = hipMemcpy(snapshots_out, h->d_snap, n * sizeof(float), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return h->hip_fail(e, "download of snapshots");
        } else {
            std::vector<uint16_t> codes(n);
            std::vector<float> maxv(h->nsnap);
            hipError_t e = hipMemcpy(codes.data(), h->d_snap16, n * sizeof(uint16_t), hipMemcpyDeviceToHost);
            if (e == hipSuccess)
                e = hipMemcpy(maxv.data(), h->d_snap_max, h->nsnap * sizeof(float), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return h->hip_fail(e, "download of snapshots");
            const auto fmt = static_cast<rtm_hip::SnapFmt>(h->snap_fmt);
            for (int s16 = 0; s16 < h->nsnap; ++s16) {
                const int k = rtm_hip::snapshot_scale_exponent(maxv[s16]);
                const size_t base = (size_t)s16 * h->n_int;
                for (size_t j = 0; j < h->n_int; ++j)
                    snapshots_out[base + j] = rtm_hip::decode_snapshot(codes[base + j], k, fmt);
            }
        }
    }
    h->times.d2h += d2h.elapsed();
    return RTM_HIP_OK;
}

int backward_impl(rtm_hip_engine* h, const rtm_hip_shot* s) {
    if (int rc = validate_shot(h, s, true)) return rc;
    rtm::Timer wall;
    if (int rc = autotune(h, Dir::Backward)) return rc;
    const rtm_hip::StepGeometry g = h->geometry(1);
    if (int rc = build_entries(h, 1, s, false, true, g)) return rc;
    if (s->nrec > 0) {
        const size_t n = (size_t)s->nrec * s->nt;
        if (int rc = h->grow(&h->d_traces, &h->traces_cap, n, "trace buffer")) return rc;
        rtm::Timer t;
        hipError_t e = hipMemcpyAsync(h->d_traces, s->traces, n * sizeof(float), hipMemcpyHostToDevice,
                                      h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "upload of traces");
        const int nent = (int)h->h_ent[1].size();
        if (int rc = h->grow(&h->d_amps, &h->amps_cap, (size_t)h->nt * nent, "amplitude table")) return rc;
        e = rtm_hip::launch_gather_amps(h->d_ent[1], nent, h->d_traces, s->nt, h->nt, h->d_amps, h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "amplitude table");
        e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) return h->hip_fail(e, "upload of traces");
        h->times.h2d += t.elapsed();
    }
    if (int rc = run_loop(h, Dir::Backward, s, false, g)) return rc;
    h->times.backward += wall.elapsed();
    return RTM_HIP_OK;
}

}  // namespace

// ============================================================================
// C ABI
// ============================================================================
extern "C" {

int rtm_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

int rtm_hip_create(int device, rtm_hip_engine** out) {
    if (!out) {
        g_null_handle_error = "null output pointer";
        return RTM_HIP_ERR_ARG;
    }
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        g_null_handle_error = "no HIP device available";
        return RTM_HIP_ERR_DEVICE;
    }
    if (device < 0 || device >= n) {
        g_null_handle_error = "device index out of range";
        return RTM_HIP_ERR_ARG;
    }
    e = hipSetDevice(device);
    if (e != hipSuccess) {
        g_null_handle_error = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return RTM_HIP_ERR_DEVICE;
    }
    auto* h = new (std::nothrow) rtm_hip_engine();
    if (!h) {
        g_null_handle_error = "out of host memory";
        return RTM_HIP_ERR_MEMORY;
    }
    h->device = device;
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        g_null_handle_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        delete h;
        return RTM_HIP_ERR_DEVICE;
    }
    // this device's copy of the no-allocation scratch (device is current)
    e = hipGetSymbolAddress(reinterpret_cast<void**>(&h->d_scratch), HIP_SYMBOL(scratch));
    if (e != hipSuccess || !h->d_scratch) {
        g_null_handle_error = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        (void)hipStreamDestroy(h->stream);
        delete h;
        return RTM_HIP_ERR_DEVICE;
    }
    *out = h;
    return RTM_HIP_OK;
}

void rtm_hip_destroy(rtm_hip_engine* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    h->release();
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
}

const char* rtm_hip_last_error(const rtm_hip_engine* h) {
    return h ? h->err.c_str() : g_null_handle_error.c_str();
}

int rtm_hip_setup(rtm_hip_engine* h, const rtm_hip_grid* g, const float* v, const rtm_hip_params* p,
                  int nt, float dt) {
    if (int rc = check_handle(h, false)) return rc;
    if (!g || !v || !p) return h->fail(RTM_HIP_ERR_ARG, "null grid, velocity or params");
    // validation of CPUReferenceRTM::setup (src/rtm/rtm_cpu.cpp:15-22)
    const int order = p->order;
    if (order != 2 && order != 4 && order != 6 && order != 8 && order != 10 && order != 12)
        return h->fail(RTM_HIP_ERR_ARG, "unsupported spatial order (use 2, 4, 6, 8, 10 or 12)");
    const int half = order / 2;
    if (p->nb < half + 1) return h->fail(RTM_HIP_ERR_ARG, "--nb must be at least order/2 + 1");
    if (g->nx <= 0 || g->nz <= 0) return h->fail(RTM_HIP_ERR_ARG, "velocity model size does not match nx*nz");
    if (p->store_interval < 1) return h->fail(RTM_HIP_ERR_ARG, "--store-interval must be >= 1");
    if (nt < 1) return h->fail(RTM_HIP_ERR_ARG, "nt must be >= 1");
    if (!(dt > 0.0f)) return h->fail(RTM_HIP_ERR_ARG, "dt must be > 0");
    if (!(g->dx > 0.0f) || !(g->dz > 0.0f)) return h->fail(RTM_HIP_ERR_ARG, "dx and dz must be > 0");
    if (p->snapshot_format != RTM_HIP_SNAP_F32 && p->snapshot_format != RTM_HIP_SNAP_F16 &&
        p->snapshot_format != RTM_HIP_SNAP_BF16)
        return h->fail(RTM_HIP_ERR_ARG, "unknown snapshot format %d", p->snapshot_format);
    if (p->image_stack != RTM_HIP_STACK_F32 && p->image_stack != RTM_HIP_STACK_F64)
        return h->fail(RTM_HIP_ERR_ARG, "unknown image stack precision %d", p->image_stack);
    const long long nxe = (long long)g->nx + 2LL * p->nb, nze = (long long)g->nz + 2LL * p->nb;
    const long long ldz = round_up((int)nze, rtm_hip::kTileZ);
    if (nxe * ldz >= (1LL << 31)) return h->fail(RTM_HIP_ERR_ARG, "extended grid exceeds 2^31 cells");
    if (!rtm_hip::step_grid_ok((int)nxe, (int)nze, (int)ldz, g->nx, g->nz))
        return h->fail(RTM_HIP_ERR_ARG, "grid too large for the step kernels (a field or the image exceeds 2 GiB)");

    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return h->hip_fail(e, "hipSetDevice");
    (void)hipStreamSynchronize(h->stream);
    h->release();

    h->nx = g->nx; h->nz = g->nz; h->dx = g->dx; h->dz = g->dz; h->ox = g->ox; h->oz = g->oz;
    h->nb = p->nb; h->order = order; h->half = half;
    h->nxe = (int)nxe; h->nze = (int)nze; h->ldz = (int)ldz;
    h->nt = nt; h->dt = dt; h->store = p->store_interval; h->verbose = p->verbose;
    h->snap_fmt = p->snapshot_format;
    h->image_stack = p->image_stack;
    h->nsnap = (nt - 1) / p->store_interval + 1;
    h->n_int = (size_t)g->nx * g->nz;
    h->n_fld = (size_t)(nxe + rtm_hip::kGuardCols) * ldz;   // guard columns of zeros (kGuardCols)
    h->snap_bytes = (size_t)h->nsnap * h->n_int *
                    (h->snap_fmt == RTM_HIP_SNAP_F32 ? sizeof(float) : sizeof(uint16_t));
    h->times = rtm_hip_times{};
    h->tuned[0] = h->tuned[1] = false;

    rtm_hip::HostTables t;
    try {
        t = rtm_hip::build_tables(g->nx, g->nz, g->dx, g->dz, p->nb, order, p->sponge_alpha, p->f0, nt, dt,
                                  v, h->ldz);
    } catch (const std::exception& ex) {
        return h->fail(RTM_HIP_ERR_ARG, "%s", ex.what());
    }
    h->coef = t.coef;

    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess) {
        const size_t need = h->snap_bytes + 6 * h->n_fld * sizeof(float) + 2 * h->n_int * sizeof(float) +
                            (h->image_stack ? 2 * h->n_int * sizeof(double) : 0);
        if (need > free_b)
            return h->fail(RTM_HIP_ERR_MEMORY,
                           "engine needs %.2f GiB of device memory (snapshots %.2f GiB), %.2f GiB free; "
                           "raise --store-interval",
                           need / 1073741824.0, h->snap_bytes / 1073741824.0, free_b / 1073741824.0);
    }

    int rc = RTM_HIP_OK;
    if (!rc) rc = h->alloc(&h->d_vdt2, h->n_fld, "vdt2");
    if (!rc) rc = h->alloc(&h->d_wx, (size_t)h->nxe + rtm_hip::kGuardCols, "wx");
    if (!rc) rc = h->alloc(&h->d_wz, (size_t)h->ldz, "wz");
    if (!rc) rc = h->alloc(&h->d_wavelet, (size_t)nt, "wavelet");
    for (int b = 0; b < 2 && !rc; ++b) rc = h->alloc(&h->d_fld[b], h->n_fld, "wavefield");
    for (int b = 0; b < 2 && !rc; ++b) rc = h->alloc(&h->d_hold[b], h->n_fld, "held wavefield");
    if (h->snap_fmt == RTM_HIP_SNAP_F32) {
        if (!rc) rc = h->alloc(&h->d_snap, h->nsnap * h->n_int, "snapshot store");
    } else {
        // one spare word: imaging reads the aligned 32-bit word holding a code
        if (!rc) rc = h->alloc(&h->d_snap16, h->nsnap * h->n_int + 2, "snapshot store");
        if (!rc) rc = h->alloc(&h->d_snap_max, (size_t)h->nsnap, "snapshot scales");
        // per-wave maxima: as many as the largest launch of any step-kernel mapping
        int parts = 0;
        for (auto f : {rtm_hip::Family::kLdsTile, rtm_hip::Family::kMarch, rtm_hip::Family::kMarchDpp,
                       rtm_hip::Family::kStream})
            for (int tx : rtm_hip::tile_widths(f)) {
                const rtm_hip::StepGeometry g = rtm_hip::step_geometry(f, h->half, tx, h->nxe, h->nze, h->ldz);
                parts = std::max(parts, f == rtm_hip::Family::kLdsTile ? g.ntiles * rtm_hip::kWaves : g.ntiles);
            }
        if (!rc) rc = h->alloc(&h->d_wave_max, 2 * (size_t)parts, "snapshot wave maxima");
        if (!rc) rc = h->alloc(&h->d_slice_max, 2 * (size_t)rtm_hip::kReduceWaves, "snapshot slice maxima");
        if (!rc) rc = h->alloc(&h->d_part_counts, 4, "snapshot maxima counts");
        if (!rc && hipMemset(h->d_part_counts, 0, 4 * sizeof(int)) != hipSuccess)
            rc = h->fail(RTM_HIP_ERR_DEVICE, "snapshot maxima counts: clearing failed");
        h->wave_max_cap = parts;
    }
    if (h->image_stack) {
        if (!rc) rc = h->alloc(&h->d_stack_i, h->n_int, "FP64 image stack");
        if (!rc) rc = h->alloc(&h->d_stack_s, h->n_int, "FP64 illumination stack");
    }
    if (!rc) rc = h->alloc(&h->d_image, h->n_int, "image");
    if (!rc) rc = h->alloc(&h->d_illum, h->n_int, "illumination");
    if (rc) {
        h->release();
        return rc;
    }
    rtm::Timer h2d;
    e = hipMemcpy(h->d_vdt2, t.vdt2.data(), t.vdt2.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_wx, t.wx.data(), t.wx.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(h->d_wz, t.wz.data(), t.wz.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(h->d_wavelet, t.wavelet.data(), t.wavelet.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(h->d_image, 0, h->n_int * sizeof(float));
    if (e == hipSuccess) e = hipMemset(h->d_illum, 0, h->n_int * sizeof(float));
    if (e == hipSuccess && h->d_snap) e = hipMemset(h->d_snap, 0, h->snap_bytes);
    if (e == hipSuccess && h->d_snap16) e = hipMemset(h->d_snap16, 0, h->snap_bytes);
    if (e == hipSuccess && h->d_snap_max) e = hipMemset(h->d_snap_max, 0, (size_t)h->nsnap * sizeof(float));
    if (e == hipSuccess && h->d_stack_i) e = hipMemset(h->d_stack_i, 0, h->n_int * sizeof(double));
    if (e == hipSuccess && h->d_stack_s) e = hipMemset(h->d_stack_s, 0, h->n_int * sizeof(double));
    for (int b = 0; b < 2 && e == hipSuccess; ++b) e = hipMemset(h->d_fld[b], 0, h->n_fld * sizeof(float));
    // a step writes output cells only: the outer ring and the guard columns
    // of a hold buffer are read as p^n / p^{n-1} and must stay zero
    for (int b = 0; b < 2 && e == hipSuccess; ++b) e = hipMemset(h->d_hold[b], 0, h->n_fld * sizeof(float));
    if (e != hipSuccess) {
        h->release();
        return h->hip_fail(e, "table upload");
    }
    h->times.h2d += h2d.elapsed();
    h->ready = true;
    return RTM_HIP_OK;
}

int rtm_hip_get_info(const rtm_hip_engine* h, rtm_hip_info* out) {
    if (int rc = check_handle(h, false)) return rc;
    if (!out) return const_cast<rtm_hip_engine*>(h)->fail(RTM_HIP_ERR_ARG, "null info");
    std::memset(out, 0, sizeof *out);
    out->nx = h->nx; out->nz = h->nz; out->nb = h->nb; out->half = h->half;
    out->nxe = h->nxe; out->nze = h->nze; out->ldz = h->ldz;
    out->nt = h->nt; out->store_interval = h->store; out->num_snapshots = h->nsnap;
    out->snapshot_bytes = h->snap_bytes;
    out->device_bytes = h->device_bytes;
    out->snapshot_format = h->snap_fmt;
    out->image_stack = h->image_stack;
    if (h->ready) {
        for (int k = 0; k < 2; ++k) {
            const rtm_hip::StepGeometry g = h->geometry(k);
            out->tile_x[k] = g.tile_x;
            out->step_family[k] = static_cast<int>(g.family);
            out->autotuned[k] = h->tuned[k] && h->family_user < 0 && h->tile_x_user == 0;
            out->waves_per_cu[k] = g.blocks_per_cu * rtm_hip::kWaves;
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, h->device) == hipSuccess) {
        std::snprintf(out->device_name, sizeof out->device_name, "%s (%s)",
                      prop.name[0] ? prop.name : "AMD Instinct", prop.gcnArchName);
    }
    return RTM_HIP_OK;
}

int rtm_hip_forward(rtm_hip_engine* h, const rtm_hip_shot* shot, float* snapshots_out, float* recorded) {
    if (int rc = check_handle(h, true)) return rc;
    (void)hipSetDevice(h->device);
    return forward_impl(h, shot, snapshots_out, recorded);
}

int rtm_hip_backward(rtm_hip_engine* h, const rtm_hip_shot* shot) {
    if (int rc = check_handle(h, true)) return rc;
    (void)hipSetDevice(h->device);
    return backward_impl(h, shot);
}

int rtm_hip_image_zero(rtm_hip_engine* h) {
    if (int rc = check_handle(h, true)) return rc;
    (void)hipSetDevice(h->device);
    hipError_t e = hipMemsetAsync(h->d_image, 0, h->n_int * sizeof(float), h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->d_illum, 0, h->n_int * sizeof(float), h->stream);
    if (e == hipSuccess && h->d_stack_i) e = hipMemsetAsync(h->d_stack_i, 0, h->n_int * sizeof(double), h->stream);
    if (e == hipSuccess && h->d_stack_s) e = hipMemsetAsync(h->d_stack_s, 0, h->n_int * sizeof(double), h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? RTM_HIP_OK : h->hip_fail(e, "image reset");
}

int rtm_hip_image_get(rtm_hip_engine* h, float* image, float* illum) {
    if (int rc = check_handle(h, true)) return rc;
    (void)hipSetDevice(h->device);
    rtm::Timer t;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (h->image_stack) {
        // the FP64 stacks, rounded once to float
        std::vector<double> tmp(h->n_int);
        if (e == hipSuccess && image) {
            e = hipMemcpy(tmp.data(), h->d_stack_i, h->n_int * sizeof(double), hipMemcpyDeviceToHost);
            for (size_t j = 0; e == hipSuccess && j < h->n_int; ++j) image[j] = (float)tmp[j];
        }
        if (e == hipSuccess && illum) {
            e = hipMemcpy(tmp.data(), h->d_stack_s, h->n_int * sizeof(double), hipMemcpyDeviceToHost);
            for (size_t j = 0; e == hipSuccess && j < h->n_int; ++j) illum[j] = (float)tmp[j];
        }
    } else {
        if (e == hipSuccess && image)
            e = hipMemcpy(image, h->d_image, h->n_int * sizeof(float), hipMemcpyDeviceToHost);
        if (e == hipSuccess && illum)
            e = hipMemcpy(illum, h->d_illum, h->n_int * sizeof(float), hipMemcpyDeviceToHost);
    }
    h->times.d2h += t.elapsed();
    return e == hipSuccess ? RTM_HIP_OK : h->hip_fail(e, "image download");
}

int rtm_hip_image_get_f64(rtm_hip_engine* h, double* image, double* illum) {
    if (int rc = check_handle(h, true)) return rc;
    if (!h->image_stack) return h->fail(RTM_HIP_ERR_ARG, "engine was set up with an FP32 image stack");
    (void)hipSetDevice(h->device);
    rtm::Timer t;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess && image)
        e = hipMemcpy(image, h->d_stack_i, h->n_int * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && illum)
        e = hipMemcpy(illum, h->d_stack_s, h->n_int * sizeof(double), hipMemcpyDeviceToHost);
    h->times.d2h += t.elapsed();
    return e == hipSuccess ? RTM_HIP_OK : h->hip_fail(e, "image download");
}

int rtm_hip_image_postprocess(rtm_hip_engine* h, int flags, float eps_relative, float* out) {
    if (int rc = check_handle(h, true)) return rc;
    if (!out) return h->fail(RTM_HIP_ERR_ARG, "null post-processing output");
    if (flags <= 0 || (flags & ~(RTM_HIP_POST_ILLUM | RTM_HIP_POST_LAPLACIAN)))
        return h->fail(RTM_HIP_ERR_ARG, "post-processing flags %d: expected RTM_HIP_POST_ILLUM and/or "
                                        "RTM_HIP_POST_LAPLACIAN", flags);
    if (h->nx > 65535) return h->fail(RTM_HIP_ERR_ARG, "post-processing supports nx <= 65535 (nx = %d)", h->nx);
    (void)hipSetDevice(h->device);
    const size_t parts = rtm_hip::post_scratch_floats(h->n_int);
    if (int rc = h->grow(&h->d_post, &h->post_cap, parts + h->n_int, "post-processing buffer")) return rc;
    float* d_out = h->d_post + parts;
    // the reference forms the inverse squared spacings on the host in float
    // (src/rtm/imaging.cpp:65-66)
    const float idx2 = 1.0f / (h->dx * h->dx);
    const float idz2 = 1.0f / (h->dz * h->dz);
    hipError_t e = rtm_hip::launch_image_post(
        h->image_stack ? nullptr : h->d_image, h->image_stack ? nullptr : h->d_illum,
        h->image_stack ? h->d_stack_i : nullptr, h->image_stack ? h->d_stack_s : nullptr, h->nx, h->nz, idx2, idz2,
        flags, eps_relative, h->d_post, d_out, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, h->n_int * sizeof(float), hipMemcpyDeviceToHost);
    return e == hipSuccess ? RTM_HIP_OK : h->hip_fail(e, "image post-processing");
}

// The same post-processing with no allocation: the partial maxima go to the
// file-scope scratch of this device, the result to the caller's device
// buffer; no hipMalloc, hipFree or host copy.
int rtm_hip_image_postprocess_noalloc(rtm_hip_engine* h, int flags, float eps_relative, float* d_out) {
    if (int rc = check_handle(h, true)) return rc;
    if (!d_out) return h->fail(RTM_HIP_ERR_ARG, "null post-processing output");
    if (flags <= 0 || (flags & ~(RTM_HIP_POST_ILLUM | RTM_HIP_POST_LAPLACIAN)))
        return h->fail(RTM_HIP_ERR_ARG, "post-processing flags %d: expected RTM_HIP_POST_ILLUM and/or "
                                        "RTM_HIP_POST_LAPLACIAN", flags);
    if (h->nx > 65535) return h->fail(RTM_HIP_ERR_ARG, "post-processing supports nx <= 65535 (nx = %d)", h->nx);
    const size_t need = rtm_hip::post_scratch_floats(h->n_int);
    if (need > (size_t)kNoAllocScratchFloats)
        return h->fail(RTM_HIP_ERR_ARG, "post-processing of %zu cells needs %zu scratch floats, more than the %d of "
                                        "the no-allocation scratch", h->n_int, need, kNoAllocScratchFloats);
    if (!h->d_scratch) return h->fail(RTM_HIP_ERR_DEVICE, "no-allocation scratch is not resolved on this device");
    hipError_t e = hipSetDevice(h->device);
    const float idx2 = 1.0f / (h->dx * h->dx);
    const float idz2 = 1.0f / (h->dz * h->dz);
    if (e == hipSuccess)
        e = rtm_hip::launch_image_post(
            h->image_stack ? nullptr : h->d_image, h->image_stack ? nullptr : h->d_illum,
            h->image_stack ? h->d_stack_i : nullptr, h->image_stack ? h->d_stack_s : nullptr, h->nx, h->nz, idx2,
            idz2, flags, eps_relative, h->d_scratch, d_out, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    return e == hipSuccess ? RTM_HIP_OK : h->hip_fail(e, "image post-processing (no allocation)");
}

int rtm_hip_migrate(rtm_hip_engine* h, int nshots, const rtm_hip_shot* shots, float* image, float* illum) {
    if (int rc = check_handle(h, true)) return rc;
    if (nshots < 0 || (nshots > 0 && !shots)) return h->fail(RTM_HIP_ERR_ARG, "bad shot list");
    rtm::Timer total;
    if (int rc = rtm_hip_image_zero(h)) return rc;
    for (int is = 0; is < nshots; ++is) {
        if (h->verbose) {
            std::printf("  shot %2d / %2d   sx=%8.1f m  sz=%6.1f m  nrec=%d\n", is + 1, nshots, shots[is].sx,
                        shots[is].sz, shots[is].nrec);
            std::fflush(stdout);
        }
        if (int rc = forward_impl(h, &shots[is], nullptr, nullptr)) return rc;
        if (int rc = backward_impl(h, &shots[is])) return rc;
    }
    if (int rc = rtm_hip_image_get(h, image, illum)) return rc;
    h->times.total += total.elapsed();
    return RTM_HIP_OK;
}

int rtm_hip_model_shot(rtm_hip_engine* h, const rtm_hip_shot* shot, float* traces_out) {
    if (int rc = check_handle(h, true)) return rc;
    if (!traces_out) return h->fail(RTM_HIP_ERR_ARG, "null trace output");
    (void)hipSetDevice(h->device);
    if (shot && shot->nrec == 0) return RTM_HIP_OK;
    return forward_impl(h, shot, nullptr, traces_out, false);
}

int rtm_hip_get_times(const rtm_hip_engine* h, rtm_hip_times* out) {
    if (int rc = check_handle(h, false)) return rc;
    if (out) *out = h->times;
    return RTM_HIP_OK;
}

int rtm_hip_reset_times(rtm_hip_engine* h) {
    if (int rc = check_handle(h, false)) return rc;
    h->times = rtm_hip_times{};
    h->launches[0] = h->launches[1] = 0;
    h->sampled[0] = h->sampled[1] = 0;
    h->sampled_ms[0] = h->sampled_ms[1] = 0.0;
    return RTM_HIP_OK;
}

int rtm_hip_set_kernel_timing(rtm_hip_engine* h, int every) {
    if (int rc = check_handle(h, false)) return rc;
    h->timing_every = every;
    return RTM_HIP_OK;
}

int rtm_hip_get_kernel_stats(const rtm_hip_engine* h, rtm_hip_kernel_stats* out) {
    if (int rc = check_handle(h, false)) return rc;
    if (!out) return RTM_HIP_OK;
    for (int k = 0; k < 2; ++k) {
        out->launches[k] = h->launches[k];
        out->sampled[k] = h->sampled[k];
        out->sampled_ms[k] = h->sampled_ms[k];
    }
    return RTM_HIP_OK;
}

int rtm_hip_set_step_kernel(rtm_hip_engine* h, int family, int tile_x) {
    if (int rc = check_handle(h, false)) return rc;
    if (family < -1 || family > RTM_HIP_STEP_STREAM)
        return h->fail(RTM_HIP_ERR_ARG, "step kernel family must be -1 or %d..%d", RTM_HIP_STEP_LDS_TILE,
                       RTM_HIP_STEP_STREAM);
    if (tile_x != 0) {
        bool ok = false;
        if (family < 0) {
            ok = tile_x == 4 || tile_x == 8 || tile_x == 16;
        } else {
            for (int t : rtm_hip::tile_widths(static_cast<rtm_hip::Family>(family))) ok |= t == tile_x;
        }
        if (!ok)
            return h->fail(RTM_HIP_ERR_ARG, "tile width %d is not available for step kernel family %d "
                           "(march-dpp: 2, 3, 4, 5, 6, 8, 16; march, lds-tile: 4, 8, 16; stream: 8, 16, 32)",
                           tile_x, family);
    }
    h->family_user = family;
    h->tile_x_user = tile_x;
    return RTM_HIP_OK;
}

}  // extern "C"
