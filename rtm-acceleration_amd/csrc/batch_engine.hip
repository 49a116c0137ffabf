// batch_engine.hip -- the batched and served engine (C ABI in rtm_hip.h,
// rtm_hip_batch_*): many shot-migration jobs, each on its own grid, advance
// together one time step per tick (batch_kernels.hip).
//
// Models.  A model is what RTMEngine::setup binds (velocity grid, RTMParams,
// time axis): its tables ((v dt)^2, tapers, wavelet, stencil coefficients)
// are built on the host with the reference's float formulas and uploaded
// once.  A registered model stays resident on the device until it is removed
// and serves any number of shots; a job submitted with its own model gets a
// private one that lives as long as the job.
//
// Scheduling.  Submitted jobs wait in a FIFO queue.  The head job is admitted
// when its device footprint (two wavefields updated in place, a scratch field
// for timing steps, two held fields, snapshot store, image, traces,
// injection tables, plus its model's tables when they are not resident yet)
// fits in what the memory budget has left and the active-job cap allows;
// admission allocates and uploads it and the launch tables are rebuilt.  The
// engine then runs ticks until the next job finishes (the host knows every
// job's clock: one step per tick), takes the finished job's image (and adds
// it into the FP64 stack when the job has a stack position), frees its
// memory and admits from the queue again: jobs that arrive while others are
// on the device start as soon as memory frees up, and their steps overlap
// with the running jobs' steps in the same launches (continuous batching).
// When a job is about to finish and the queue head fits in the memory free
// now, the head is admitted ahead, while the device runs the ticks up to
// that finish; it joins the active set at the change.
//
// Launches.  Jobs are grouped by stencil half-width and snapshot width; each
// group has one launch per tick over the concatenation of its jobs' tiles,
// with the tile width and cross-lane method picked per group and load class
// (octave of the group's extended cells, and which kernels run) by timing dry
// steps on the live jobs.  Runs of up to 32 ticks (the group launches on
// concurrent streams, one clock update) are captured in hipGraphs and
// replayed while ticks are small enough to be launch-bound.  A job alone, or
// two or three jobs of which one fills the chip, run on the per-shot step
// kernels instead, each on a stream of its own.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "arena.hpp"
#include "batch_args.hpp"
#include "batch_kernels.hpp"
#include "host_tables.hpp"
#include "rtm_hip.h"

using rtm_hip::Entry;
using rtm_hip::JobDesc;

namespace {

thread_local std::string g_batch_null_error;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int round_up(int v, int m) { return (v + m - 1) / m * m; }

// One set-up model: grid, parameters, time axis and (when resident) its
// device tables.
struct Model {
    int id = 0;
    bool registered = false;         // stays resident until removed
    bool removing = false;           // removal requested while jobs still use it
    int refs = 0;                    // queued + active jobs on it
    rtm_hip_grid grid{};
    rtm_hip_params params{};
    int nt = 0;
    float dt = 0.0f;
    std::vector<float> v;            // host velocity until the tables are uploaded
    int half = 0, nxe = 0, nze = 0, ldz = 0, nsnap = 0;
    size_t n_int = 0, n_fld = 0, bytes = 0;
    rtm_hip::Coeffs coef{};
    bool resident = false;
    float *vdt2 = nullptr, *wx = nullptr, *wz = nullptr, *wavelet = nullptr;
    float* zero = nullptr;        // a field of zeros: p^{n-1}, p^n of a phase's first steps
};

// One submitted shot migration and, once admitted, its device buffers.
struct Job {
    long long id = 0;
    int state = RTM_HIP_JOB_QUEUED;
    std::shared_ptr<Model> model;
    float sx = 0.0f, sz = 0.0f;
    std::vector<float> rx, rz, traces;
    int trace_nt = 0;
    float* d_traces = nullptr;       // device copy of the traces (job allocation)
    int stack_ix0 = -1, stack_iz0 = -1;
    size_t bytes = 0;                // device bytes of the job itself (model excluded)
    std::vector<void*> allocs;       // freed when the job retires
    std::vector<void*> entry_allocs; // injection tables of the current tile width
    JobDesc desc{};
    int tile_x_built = 0;            // tile width the injection tables were built for
                                     // (negative: entry-free tables for timing, stub_job_entries)
    int steps_left = 0;              // ticks until done (host mirror of the device clock)
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<float> image, illum;
};

// Kernel choice of a half-width group.
struct Tuning {
    rtm_hip::Family family = rtm_hip::Family::kMarchDpp;
    int tile_x = 4;
    int blocks_per_cu = 0;           // occupancy cap (0 = none) and the LDS that enforces it
    int lds_bytes = 0;
};

// The best kernel depends on how much work one launch holds: a lone small
// grid wants narrow tiles (waves to fill 256 CUs), many large grids want wide
// ones (less halo re-read).  Loads are classed by octave of the group's
// extended cells, and each class is tuned once, on first sight.
struct Group : Tuning {
    bool tuned = false;
    int load_class = -1;             // class of the job set the tables were built for
    std::map<int, Tuning> by_load;   // tuned choice per load class
    int ntiles = 0;
    int* d_tile_job = nullptr;
    size_t cap = 0;
};

int load_class_of(double ext_cells) { return (int)std::floor(std::log2(std::max(ext_cells, 1.0))); }

// Jobs share a launch when they have the same stencil half-width and
// snapshot width (the FP32 and 16-bit kernels are separate builds).
int group_key(const Model& m) { return 2 * m.half + (m.params.snapshot_format != RTM_HIP_SNAP_F32 ? 1 : 0); }
int key_half(int key) { return key / 2; }
bool key_snap16(int key) { return (key & 1) != 0; }

}  // namespace

// Streams a run of ticks spreads its launch groups over: groups are
// independent within a run (each job reads only its own state and the run's
// tick arguments), so their launches run concurrently and fill each other's
// tails, where launches on one stream would each drain the device.
constexpr int kGroupStreams = 3;

struct rtm_hip_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t group_streams[kGroupStreams - 1] = {};   // groups 1, 2, ... of a run (group 0: stream)
    hipEvent_t fork_ev = nullptr;                       // the run's tick arguments are ready
    // the ticks a rtm_hip_batch_run call with jobs left running leaves on the
    // device (the next such call waits for them)
    hipEvent_t chunk_ev[2] = {};
    int chunk_cur = 0;
    bool chunk_pending = false;
    hipEvent_t join_ev[kGroupStreams - 1] = {};         // a group stream's launches of the run are done
    std::string err;
    rtm_hip_batch_config cfg{};
    size_t budget = 0, used = 0;
    long long next_job = 1;
    int next_model = 1;
    std::map<int, std::shared_ptr<Model>> models;    // registered
    std::map<long long, std::unique_ptr<Job>> jobs;
    std::deque<Job*> queue;
    std::vector<Job*> active;
    // jobs admitted ahead (stage_ahead): device state ready, joining the
    // active set at its next change, before anything still in the queue
    std::vector<Job*> staged;
    JobDesc* d_jobs = nullptr;
    rtm_hip::TickArgs* d_targs = nullptr;            // [jobs_cap][kRunTicks] (launch_tick_args)
    // while the autotuner times candidates, the launch tables give every job
    // entry-free tile tables (stub_job_entries) in this table of zeros
    bool timing = false;
    int* d_zero_ptr = nullptr;
    size_t zero_cap = 0;
    size_t jobs_cap = 0;
    std::map<int, Group> groups;                     // by group_key (half-width, snapshot width)
    // One launch per tick for every job (cfg.merge_groups, with jobs of more
    // than one group and the DPP march family): the tiles of all jobs in one
    // table, the kernel picking each wave's tile build (launch_batch_merged).
    Group merged;
    bool merged_on = false;
    bool dirty = true;
    // captured runs of ticks, by the run's launch pattern: its length and,
    // per tick and group, the lean or full build (a few patterns recur while
    // the job set stays the same)
    struct Captured {
        std::string pattern;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    std::vector<Captured> graphs;
    int stack_nx = 0, stack_nz = 0;
    double* d_stack_i = nullptr;
    double* d_stack_s = nullptr;
    rtm_hip::Arena arena;                            // reserved job memory, if any
    rtm_hip_batch_stats stats{};

    // Job and model buffers: from the reserved arena when there is one.
    hipError_t dev_alloc(void** p, size_t bytes) {
        if (arena.active()) {
            *p = arena.alloc(bytes);
            return *p ? hipSuccess : hipErrorOutOfMemory;
        }
        return hipMalloc(p, std::max<size_t>(bytes, 1));
    }
    void dev_free(void* p) {
        if (!p) return;
        if (arena.contains(p)) arena.release(p);
        else (void)hipFree(p);
    }

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
        return fail(e == hipErrorOutOfMemory ? RTM_HIP_ERR_MEMORY : RTM_HIP_ERR_DEVICE, "%s: %s", what,
                    hipGetErrorString(e));
    }
    void drop_graph() {
        for (Captured& c : graphs) {
            if (c.exec) (void)hipGraphExecDestroy(c.exec);
            if (c.graph) (void)hipGraphDestroy(c.graph);
        }
        graphs.clear();
    }
};

namespace {

constexpr double kGiB = 1073741824.0;

// ---- models ---------------------------------------------------------------

// Validates the setup arguments (the checks and messages of
// CPUReferenceRTM::setup, src/rtm/rtm_cpu.cpp:9-31) and derives the sizes.
int model_init(rtm_hip_batch* b, Model& m, const rtm_hip_grid* grid, const float* v, const rtm_hip_params* params,
               int nt, float dt) {
    if (!grid || !v || !params) return b->fail(RTM_HIP_ERR_ARG, "null model argument");
    m.grid = *grid;
    m.params = *params;
    m.nt = nt;
    m.dt = dt;
    const int order = m.params.order;
    if (order != 2 && order != 4 && order != 6 && order != 8 && order != 10 && order != 12)
        return b->fail(RTM_HIP_ERR_ARG, "unsupported spatial order (use 2, 4, 6, 8, 10 or 12)");
    m.half = order / 2;
    if (m.params.nb < m.half + 1) return b->fail(RTM_HIP_ERR_ARG, "--nb must be at least order/2 + 1");
    if (m.grid.nx <= 0 || m.grid.nz <= 0) return b->fail(RTM_HIP_ERR_ARG, "velocity model size does not match nx*nz");
    if (m.params.store_interval < 1) return b->fail(RTM_HIP_ERR_ARG, "--store-interval must be >= 1");
    if (m.nt < 1 || !(m.dt > 0.0f)) return b->fail(RTM_HIP_ERR_ARG, "nt must be >= 1 and dt > 0");
    const int fmt = m.params.snapshot_format;
    if (fmt != RTM_HIP_SNAP_F32 && fmt != RTM_HIP_SNAP_F16 && fmt != RTM_HIP_SNAP_BF16)
        return b->fail(RTM_HIP_ERR_ARG, "unknown snapshot format");
    m.nxe = m.grid.nx + 2 * m.params.nb;
    m.nze = m.grid.nz + 2 * m.params.nb;
    m.ldz = round_up(m.nze, rtm_hip::kTileZ);
    if ((long long)m.nxe * m.ldz >= (1LL << 31)) return b->fail(RTM_HIP_ERR_ARG, "extended grid exceeds 2^31 cells");
    if (!rtm_hip::step_grid_ok(m.nxe, m.nze, m.ldz, m.grid.nx, m.grid.nz))
        return b->fail(RTM_HIP_ERR_ARG, "grid too large for the step kernels (a field or the image exceeds 2 GiB)");
    m.nsnap = (m.nt - 1) / m.params.store_interval + 1;
    m.n_int = (size_t)m.grid.nx * m.grid.nz;
    m.n_fld = (size_t)(m.nxe + rtm_hip::kGuardCols) * m.ldz;   // guard columns of zeros (kGuardCols)
    m.bytes = (2 * m.n_fld + (size_t)m.nxe + rtm_hip::kGuardCols + (size_t)m.ldz + (size_t)m.nt) * sizeof(float);
    m.v.assign(v, v + m.n_int);
    return RTM_HIP_OK;
}

void model_release(rtm_hip_batch* b, Model& m) {
    for (float* p : {m.vdt2, m.wx, m.wz, m.wavelet, m.zero}) b->dev_free(p);
    m.vdt2 = m.wx = m.wz = m.wavelet = m.zero = nullptr;
    m.resident = false;
}

int model_upload(rtm_hip_batch* b, Model& m) {
    rtm_hip::HostTables t;
    try {
        t = rtm_hip::build_tables(m.grid.nx, m.grid.nz, m.grid.dx, m.grid.dz, m.params.nb, m.params.order,
                                  m.params.sponge_alpha, m.params.f0, m.nt, m.dt, m.v.data(), m.ldz);
    } catch (const std::exception& ex) {
        return b->fail(RTM_HIP_ERR_ARG, "%s", ex.what());
    }
    m.coef = t.coef;
    struct Up {
        float** dst;
        const std::vector<float>* src;
    } ups[] = {{&m.vdt2, &t.vdt2}, {&m.wx, &t.wx}, {&m.wz, &t.wz}, {&m.wavelet, &t.wavelet}};
    for (const Up& u : ups) {
        hipError_t e = b->dev_alloc(reinterpret_cast<void**>(u.dst), u.src->size() * sizeof(float));
        if (e == hipSuccess)
            e = hipMemcpy(*u.dst, u.src->data(), u.src->size() * sizeof(float), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            model_release(b, m);
            return b->hip_fail(e, "model tables");
        }
    }
    hipError_t ez = b->dev_alloc(reinterpret_cast<void**>(&m.zero), m.n_fld * sizeof(float));
    if (ez == hipSuccess) ez = hipMemset(m.zero, 0, m.n_fld * sizeof(float));
    if (ez != hipSuccess) {
        model_release(b, m);
        return b->hip_fail(ez, "model zero field");
    }
    m.resident = true;
    m.v.clear();
    m.v.shrink_to_fit();
    b->used += m.bytes;
    return RTM_HIP_OK;
}

// Frees a model's tables once nothing needs them any more.
void model_maybe_free(rtm_hip_batch* b, const std::shared_ptr<Model>& m) {
    if (m->refs > 0 || !m->resident) return;
    if (m->registered && !m->removing) return;
    model_release(b, *m);
    b->used -= m->bytes;
}

// ---- jobs -----------------------------------------------------------------

// Tiles of the job's largest batched launch (the narrowest march tile).
size_t max_job_tiles(const Model& m) {
    size_t n = 0;
    for (auto f : {rtm_hip::Family::kMarch, rtm_hip::Family::kMarchDpp, rtm_hip::Family::kStream})
        for (int tx : rtm_hip::tile_widths(f))
            n = std::max(n, (size_t)rtm_hip::step_geometry(f, m.half, tx, m.nxe, m.nze, m.ldz).ntiles);
    return n;
}

size_t job_footprint(const Model& m, size_t nrec, size_t ntraces) {
    const size_t snap_elem = m.params.snapshot_format == RTM_HIP_SNAP_F32 ? 4 : 2;
    const size_t tiles_bound = (size_t)(m.nxe / 2 + 2) * (m.nze / 52 + 2) + 1;   // tile_x >= 2, strip >= 52
    size_t bytes = 0;
    bytes += 5 * m.n_fld * sizeof(float);                   // two wavefields, a dry-step output, two held
    bytes += (size_t)m.nsnap * m.n_int * snap_elem;         // snapshot store
    if (snap_elem == 2)                                     // spare word, scales and their bookkeeping
        bytes += ((size_t)m.nsnap + 2 * tiles_bound + 2 * rtm_hip::kReduceWaves + 5) * sizeof(float);
    bytes += 2 * m.n_int * sizeof(float);                   // image, illumination
    bytes += ntraces * sizeof(float);                       // receiver data
    bytes += (size_t)m.nt * nrec * sizeof(float);           // ... in backward-entry order
    bytes += 2 * (tiles_bound * sizeof(int) + (nrec + 1) * sizeof(Entry));
    return bytes;
}

int job_init(rtm_hip_batch* b, Job& j, const std::shared_ptr<Model>& m, const rtm_hip_shot* shot, int ix0, int iz0) {
    if (!shot || !shot->rx || !shot->rz || !shot->traces || shot->nrec <= 0)
        return b->fail(RTM_HIP_ERR_ARG, "a migration job needs receivers and traces");
    if (shot->nt < m->nt)
        return b->fail(RTM_HIP_ERR_ARG, "shot has %d samples per trace, fewer than nt=%d", shot->nt, m->nt);
    if (ix0 >= 0) {
        if (!b->d_stack_i) return b->fail(RTM_HIP_ERR_ARG, "job has a stack position but no stack is set");
        if (iz0 < 0 || ix0 + m->grid.nx > b->stack_nx || iz0 + m->grid.nz > b->stack_nz)
            return b->fail(RTM_HIP_ERR_ARG, "job image does not fit in the stack at its position");
    }
    j.model = m;
    j.sx = shot->sx;
    j.sz = shot->sz;
    j.rx.assign(shot->rx, shot->rx + shot->nrec);
    j.rz.assign(shot->rz, shot->rz + shot->nrec);
    j.trace_nt = shot->nt;
    j.traces.assign(shot->traces, shot->traces + (size_t)shot->nrec * shot->nt);
    j.stack_ix0 = ix0;
    j.stack_iz0 = ix0 >= 0 ? iz0 : -1;
    j.bytes = job_footprint(*m, j.rx.size(), j.traces.size());
    return RTM_HIP_OK;
}

// Device bytes admission needs for the job now.
size_t admission_bytes(const Job& j) { return j.bytes + (j.model->resident ? 0 : j.model->bytes); }

template <typename T>
int job_alloc(rtm_hip_batch* b, std::vector<void*>& list, T** p, size_t count, const char* what) {
    hipError_t e = b->dev_alloc(reinterpret_cast<void**>(p), count * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return b->hip_fail(e, what);
    }
    list.push_back(*p);
    return RTM_HIP_OK;
}

void free_entries(rtm_hip_batch* b, Job& j) {
    for (void* p : j.entry_allocs) b->dev_free(p);
    j.entry_allocs.clear();
    j.tile_x_built = 0;
}

void job_free(rtm_hip_batch* b, Job& j) {
    free_entries(b, j);
    for (void* p : j.allocs) b->dev_free(p);
    j.allocs.clear();
}

// Nearest-grid-point cell of a position, clamped to the interior
// (src/rtm/rtm_cpu.cpp:99-123).
void job_cell(const Model& m, float x, float z, int* ixe, int* ize) {
    const int ix0 = static_cast<int>(std::lround((x - m.grid.ox) / m.grid.dx));
    const int iz0 = static_cast<int>(std::lround((z - m.grid.oz) / m.grid.dz));
    *ixe = std::clamp(ix0, 0, m.grid.nx - 1) + m.params.nb;
    *ize = std::clamp(iz0, 0, m.grid.nz - 1) + m.params.nb;
}

// Per-tile injection tables of the job for tile width tile_x (forward: the
// source; backward: every receiver, in receiver order within a cell), both
// directions in one block uploaded with one copy: tile pointers, entries,
// tile pointers, entries, each part at an even int offset (8-byte aligned
// entries).
int build_job_entries(rtm_hip_batch* b, Job& j, int tile_x) {
    free_entries(b, j);
    const Model& m = *j.model;
    const rtm_hip::StepGeometry g =
        rtm_hip::step_geometry(rtm_hip::Family::kMarchDpp, m.half, tile_x, m.nxe, m.nze, m.ldz);
    std::vector<int> blob;
    size_t off_ptr[2], off_ent[2];
    for (int k = 0; k < 2; ++k) {
        struct T {
            int tile, cell, code;
        };
        std::vector<T> v;
        int ixe, ize;
        if (k == 0) {
            job_cell(m, j.sx, j.sz, &ixe, &ize);
            v.push_back({rtm_hip::tile_of_cell(g, ixe, ize), ixe * m.ldz + ize, -1});
        } else {
            for (size_t ir = 0; ir < j.rx.size(); ++ir) {
                job_cell(m, j.rx[ir], j.rz[ir], &ixe, &ize);
                v.push_back({rtm_hip::tile_of_cell(g, ixe, ize), ixe * m.ldz + ize, (int)ir});
            }
        }
        std::sort(v.begin(), v.end(), [](const T& a, const T& c) {
            if (a.tile != c.tile) return a.tile < c.tile;
            if (a.cell != c.cell) return a.cell < c.cell;
            return a.code < c.code;
        });
        off_ptr[k] = blob.size();
        blob.resize(off_ptr[k] + (size_t)g.ntiles + 1, 0);
        int* ptr = blob.data() + off_ptr[k];
        for (const T& t : v) ++ptr[(size_t)t.tile + 1];
        for (int t = 0; t < g.ntiles; ++t) ptr[t + 1] += ptr[t];
        if (blob.size() & 1) blob.push_back(0);
        off_ent[k] = blob.size();
        for (const T& t : v) {
            blob.push_back(t.cell);
            blob.push_back(t.code);
        }
    }
    int* d = nullptr;
    if (int rc = job_alloc(b, j.entry_allocs, &d, blob.size(), "injection tables")) return rc;
    hipError_t e = hipMemcpy(d, blob.data(), blob.size() * sizeof(int), hipMemcpyHostToDevice);
    if (e != hipSuccess) return b->hip_fail(e, "injection table upload");
    for (int k = 0; k < 2; ++k) {
        j.desc.tile_ptr[k] = d + off_ptr[k];
        j.desc.entries[k] = reinterpret_cast<const Entry*>(d + off_ent[k]);
    }
    // the receiver amplitudes in the new backward-entry order
    e = rtm_hip::launch_gather_amps(j.desc.entries[1], j.desc.amp_count, j.d_traces, j.trace_nt, m.nt, j.desc.amps,
                                    b->stream);
    if (e != hipSuccess) return b->hip_fail(e, "amplitude table");
    j.desc.ntz = g.ntz;
    j.desc.ntz_magic = g.ntz_magic;
    j.desc.ntx = g.ntx;
    j.desc.ntiles = g.ntiles;
    j.tile_x_built = tile_x;
    return RTM_HIP_OK;
}

// The tile geometry of width tile_x with no injection entries, for the
// autotuner's dry steps, which inject and record nothing: every tile's entry
// range is empty in the handle's shared table of zeros.  The job's own tables
// are rebuilt for the width the search picks.
int stub_job_entries(rtm_hip_batch* b, Job& j, int tile_x) {
    free_entries(b, j);
    const Model& m = *j.model;
    const rtm_hip::StepGeometry g =
        rtm_hip::step_geometry(rtm_hip::Family::kMarchDpp, m.half, tile_x, m.nxe, m.nze, m.ldz);
    const size_t need = (size_t)g.ntiles + 1;
    if (need > b->zero_cap) {
        if (b->d_zero_ptr) (void)hipFree(b->d_zero_ptr);
        b->d_zero_ptr = nullptr;
        b->zero_cap = 0;
        hipError_t e = hipMalloc(&b->d_zero_ptr, need * sizeof(int));
        if (e == hipSuccess) e = hipMemset(b->d_zero_ptr, 0, need * sizeof(int));
        if (e != hipSuccess) return b->hip_fail(e, "empty tile table");
        b->zero_cap = need;
    }
    for (int k = 0; k < 2; ++k) {
        j.desc.tile_ptr[k] = b->d_zero_ptr;
        j.desc.entries[k] = reinterpret_cast<const Entry*>(b->d_zero_ptr);
    }
    j.desc.ntz = g.ntz;
    j.desc.ntz_magic = g.ntz_magic;
    j.desc.ntx = g.ntx;
    j.desc.ntiles = g.ntiles;
    j.tile_x_built = -tile_x;
    return RTM_HIP_OK;
}

// Allocate, zero and upload everything an admitted job needs.
int admit_job(rtm_hip_batch* b, Job& j) {
    Model& m = *j.model;
    if (!m.resident)
        if (int rc = model_upload(b, m)) return rc;
    JobDesc& d = j.desc;
    d = JobDesc{};
    d.nxe = m.nxe; d.nze = m.nze; d.ldz = m.ldz; d.nb = m.params.nb; d.nx = m.grid.nx; d.nz = m.grid.nz;
    d.half = m.half; d.nt = m.nt; d.store = m.params.store_interval; d.snap_fmt = m.params.snapshot_format;
    d.coef = m.coef;
    d.n_int = (long long)m.n_int;
    d.bwd_base = m.nt & 1;                  // not the buffer of the forward phase's last field
    d.vdt2 = m.vdt2; d.wx = m.wx; d.wz = m.wz; d.wavelet = m.wavelet;
    d.zero = m.zero;
    // 16-bit jobs image their top stored step from its pending field at
    // backward step 0, so their pairing starts one stored step lower
    d.pair_first = rtm_hip::pair_first(m.nt, m.params.store_interval, d.snap_fmt != RTM_HIP_SNAP_F32);
    float *tr = nullptr, *img = nullptr, *ill = nullptr;
    int rc = job_alloc(b, j.allocs, &tr, j.traces.size(), "traces");
    float* amps = nullptr;
    if (!rc) rc = job_alloc(b, j.allocs, &amps, (size_t)m.nt * j.rx.size(), "amplitude table");
    if (!rc) rc = job_alloc(b, j.allocs, &img, m.n_int, "image");
    if (!rc) rc = job_alloc(b, j.allocs, &ill, m.n_int, "illumination");
    for (int k = 0; k < 3 && !rc; ++k) rc = job_alloc(b, j.allocs, &d.fld[k], m.n_fld, "wavefield");
    for (int k = 0; k < 2 && !rc; ++k) rc = job_alloc(b, j.allocs, &d.hold[k], m.n_fld, "held wavefield");
    if (d.snap_fmt == RTM_HIP_SNAP_F32) {
        if (!rc) rc = job_alloc(b, j.allocs, &d.snap32, (size_t)m.nsnap * m.n_int, "snapshot store");
    } else {
        // one spare word: imaging reads the aligned 32-bit word holding a code
        if (!rc) rc = job_alloc(b, j.allocs, &d.snap16, (size_t)m.nsnap * m.n_int + 2, "snapshot store");
        if (!rc) rc = job_alloc(b, j.allocs, &d.snap_max, (size_t)m.nsnap, "snapshot scales");
        d.parts_cap = (int)max_job_tiles(m);
        if (!rc) rc = job_alloc(b, j.allocs, &d.wave_max, 2 * (size_t)d.parts_cap, "snapshot tile maxima");
        if (!rc) rc = job_alloc(b, j.allocs, &d.slice_max, 2 * (size_t)rtm_hip::kReduceWaves, "snapshot slice maxima");
        if (!rc) rc = job_alloc(b, j.allocs, &d.part_counts, 4, "snapshot maxima counts");
    }
    if (rc) return rc;
    hipError_t e = hipMemcpy(tr, j.traces.data(), j.traces.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(img, 0, m.n_int * sizeof(float));
    if (e == hipSuccess) e = hipMemset(ill, 0, m.n_int * sizeof(float));
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMemset(d.fld[k], 0, m.n_fld * sizeof(float));
    // a step writes output cells only: the outer ring and the guard columns
    // of a hold buffer are read as p^n / p^{n-1} and must stay zero
    for (int k = 0; k < 2 && e == hipSuccess; ++k) e = hipMemset(d.hold[k], 0, m.n_fld * sizeof(float));
    if (e == hipSuccess && d.snap_max) e = hipMemset(d.snap_max, 0, (size_t)m.nsnap * sizeof(float));
    if (e == hipSuccess && d.part_counts) e = hipMemset(d.part_counts, 0, 4 * sizeof(int));
    if (e != hipSuccess) return b->hip_fail(e, "job upload");
    j.d_traces = tr;
    d.amps = amps;
    d.amp_count = (int)j.rx.size();
    d.image = img; d.illum = ill;
    d.steps = 2 * m.nt + (d.snap_fmt == RTM_HIP_SNAP_F32 ? 0 : 1);     // + the bookkeeping tick
    j.steps_left = d.steps;
    j.traces.clear();
    j.traces.shrink_to_fit();
    return RTM_HIP_OK;
}

// ---- launch tables, ticks, autotuning ------------------------------------

// The launch tables of every half-width group and the device descriptors.
// Ticks still running on the stream read the descriptors, the step arguments
// and the tables rewritten here, so the stream is drained first.
int rebuild_tables(rtm_hip_batch* b);

int rebuild(rtm_hip_batch* b) {
    const double t0 = now_s();
    const int rc = rebuild_tables(b);
    b->stats.rebuild_seconds += now_s() - t0;
    return rc;
}

bool run_concurrently(const rtm_hip_batch* b);

// Whether the active jobs run on the per-shot step kernels (one job, or two
// or three on streams of their own) rather than on the batch kernel.
bool per_shot_kernels(const rtm_hip_batch* b) { return b->active.size() == 1 || run_concurrently(b); }

// The tile table of one launch over the active jobs idx: the tuning of the
// launch's load class (when it has been tuned before), every job's entry
// tables at the launch's tile width, and tile_job uploaded.  The class is
// the octave of the launch's extended cells and whether its jobs run on the
// per-shot kernels, which favour other tiles than the batch kernel.
int build_group_table(rtm_hip_batch* b, Group& g, const std::vector<int>& idx) {
    if (b->cfg.kernel_family == 0) {
        double cells = 0;
        for (int k : idx) cells += (double)b->active[k]->model->nxe * b->active[k]->model->nze;
        const int cls = 2 * load_class_of(cells) + (per_shot_kernels(b) ? 1 : 0);
        if (cls != g.load_class) {
            g.load_class = cls;
            auto it = g.by_load.find(cls);
            if (it != g.by_load.end()) static_cast<Tuning&>(g) = it->second;
            g.tuned = it != g.by_load.end();
        }
    }
    std::vector<int> tile_job;
    for (int k : idx) {
        Job& j = *b->active[k];
        if (b->timing && j.tile_x_built != -g.tile_x) {
            if (int rc = stub_job_entries(b, j, g.tile_x)) return rc;
        } else if (!b->timing && j.tile_x_built != g.tile_x) {
            if (int rc = build_job_entries(b, j, g.tile_x)) return rc;
        }
        j.desc.tile_first = (int)tile_job.size();
        tile_job.insert(tile_job.end(), (size_t)j.desc.ntiles, k);
    }
    g.ntiles = (int)tile_job.size();
    if ((size_t)g.ntiles > g.cap) {
        if (g.d_tile_job) (void)hipFree(g.d_tile_job);
        g.cap = (size_t)g.ntiles;
        hipError_t e = hipMalloc(&g.d_tile_job, g.cap * sizeof(int));
        if (e != hipSuccess) {
            g.d_tile_job = nullptr;
            g.cap = 0;
            return b->hip_fail(e, "tile table");
        }
    }
    hipError_t e = hipMemcpy(g.d_tile_job, tile_job.data(), tile_job.size() * sizeof(int), hipMemcpyHostToDevice);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "tile table upload");
}

int rebuild_tables(rtm_hip_batch* b) {
    if (hipError_t e = hipStreamSynchronize(b->stream); e != hipSuccess) return b->hip_fail(e, "batch synchronisation");
    b->drop_graph();
    if (b->active.size() > b->jobs_cap) {
        if (b->d_jobs) (void)hipFree(b->d_jobs);
        if (b->d_targs) (void)hipFree(b->d_targs);
        b->d_jobs = nullptr;
        b->d_targs = nullptr;
        b->jobs_cap = std::max<size_t>(b->active.size(), 2 * b->jobs_cap);
        hipError_t e = hipMalloc(&b->d_jobs, b->jobs_cap * sizeof(JobDesc));
        if (e == hipSuccess)
            e = hipMalloc(&b->d_targs, b->jobs_cap * rtm_hip::kRunTicks * sizeof(rtm_hip::TickArgs));
        if (e != hipSuccess) {
            b->jobs_cap = 0;
            return b->hip_fail(e, "job descriptors");
        }
    }
    std::map<int, std::vector<int>> by_key;
    for (size_t k = 0; k < b->active.size(); ++k) by_key[group_key(*b->active[k]->model)].push_back((int)k);
    const int fam = b->cfg.kernel_family;
    b->merged_on = b->cfg.merge_groups && by_key.size() > 1 && (fam == 0 || fam == RTM_HIP_STEP_MARCH_DPP);
    for (auto& [key, g] : b->groups)
        if (b->merged_on || !by_key.count(key)) g.ntiles = 0;
    if (b->merged_on) {
        std::vector<int> all(b->active.size());
        for (size_t k = 0; k < all.size(); ++k) all[k] = (int)k;
        if (int rc = build_group_table(b, b->merged, all)) return rc;
    } else {
        b->merged.ntiles = 0;
        for (auto& [key, idx] : by_key)
            if (int rc = build_group_table(b, b->groups[key], idx)) return rc;
    }
    // every clock is known on the host: one step per tick since admission
    std::vector<JobDesc> descs;
    for (Job* j : b->active) {
        j->desc.done = j->desc.steps - j->steps_left;
        descs.push_back(j->desc);
    }
    if (!descs.empty()) {
        hipError_t e = hipMemcpy(b->d_jobs, descs.data(), descs.size() * sizeof(JobDesc), hipMemcpyHostToDevice);
        if (e != hipSuccess) return b->hip_fail(e, "descriptor upload");
    }
    b->dirty = false;
    b->stats.rebuilds += 1;
    return RTM_HIP_OK;
}

// Whether no job of group `key` has snapshot work `ahead` ticks after the
// host's clocks (step_is_lean on the job's step arguments, as job_args sets
// them up on the device), so the tick can run the group's lean build: no
// job stores, reduces or quantises a snapshot, and none images.  Jobs done
// by then idle, and a 16-bit job's bookkeeping tick does not run the march
// tile.  Other kernel families have no lean build.
bool jobs_tick_lean(const rtm_hip_batch* b, int key, int ahead) {
    for (const Job* j : b->active) {
        const int jkey = group_key(*j->model);
        if (key >= 0 && jkey != key) continue;
        const bool s16 = key_snap16(jkey);
        const int total = j->desc.steps - j->steps_left + ahead;
        if (total >= j->desc.steps || (s16 && total == j->desc.nt)) continue;
        const rtm_hip::Dir d = total < j->desc.nt ? rtm_hip::Dir::Forward : rtm_hip::Dir::Backward;
        const rtm_hip::StepArgs a = s16 ? rtm_hip::job_args<true>(j->desc, total, 0)
                                        : rtm_hip::job_args<false>(j->desc, total, 0);
        if (!rtm_hip::step_is_lean(d, a, s16)) return false;
    }
    return true;
}
bool group_tick_lean(const rtm_hip_batch* b, int key, const Group& g, int ahead) {
    return g.family == rtm_hip::Family::kMarchDpp && jobs_tick_lean(b, key, ahead);
}
// The merged launch picks the lean build per job on the device; for the
// launch statistics a merged tick counts as lean when every job's is.
bool merged_tick_lean(const rtm_hip_batch* b, int ahead) { return jobs_tick_lean(b, -1, ahead); }

// The launch pattern of a run of `ticks` ticks starting `ahead` ticks after
// the host's clocks: its length, then per tick and group '1' (lean) or '0'.
std::string run_pattern(const rtm_hip_batch* b, int ticks, int ahead) {
    // the merged launch is the same launch whatever its jobs' builds
    if (b->merged_on) return std::to_string(ticks) + "m";
    std::string p = std::to_string(ticks) + ":";
    for (int sub = 0; sub < ticks; ++sub)
        for (const auto& [key, g] : b->groups)
            if (g.ntiles) p += group_tick_lean(b, key, g, ahead + sub) ? '1' : '0';
    return p;
}

// A run of `ticks` ticks (at most kRunTicks) starting `ahead` ticks after the
// host's clocks: the tick arguments of the run, the group launches of every
// tick, then one clock update for the run.
int enqueue_ticks(rtm_hip_batch* b, int ticks, int ahead) {
    if (ticks < 1 || ticks > rtm_hip::kRunTicks)
        return b->fail(RTM_HIP_ERR_ARG, "internal: a run of %d ticks (1..%d)", ticks, rtm_hip::kRunTicks);
    hipError_t ea = rtm_hip::launch_tick_args(b->d_jobs, (int)b->active.size(), b->d_targs, ticks, 0, b->stream);
    if (ea != hipSuccess) return b->hip_fail(ea, "tick arguments");
    if (b->merged_on) {
        const Group& g = b->merged;
        for (int sub = 0; sub < ticks; ++sub) {
            hipError_t e = rtm_hip::launch_batch_merged(g.tile_x, b->d_jobs, b->d_targs, g.d_tile_job, g.ntiles, sub,
                                                        b->stream);
            if (e != hipSuccess) return b->hip_fail(e, "batched step launch");
        }
        hipError_t e = rtm_hip::launch_batch_advance(b->d_jobs, (int)b->active.size(), ticks, b->stream);
        return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "clock launch");
    }
    // group i of the run on stream i mod kGroupStreams (0: the handle's own)
    std::vector<std::pair<int, Group*>> run_groups;
    for (auto& [key, g] : b->groups)
        if (g.ntiles) run_groups.push_back({key, &g});
    const int nstreams = std::min<int>((int)run_groups.size(), kGroupStreams);
    auto stream_of = [&](size_t i) { return i % kGroupStreams == 0 ? b->stream : b->group_streams[i % kGroupStreams - 1]; };
    if (nstreams > 1) {
        hipError_t e = hipEventRecord(b->fork_ev, b->stream);
        for (int k = 1; k < nstreams && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->group_streams[k - 1], b->fork_ev, 0);
        if (e != hipSuccess) return b->hip_fail(e, "group stream fork");
    }
    for (size_t i = 0; i < run_groups.size(); ++i) {
        const int key = run_groups[i].first;
        const Group& g = *run_groups[i].second;
        for (int sub = 0; sub < ticks; ++sub) {
            const bool lean = group_tick_lean(b, key, g, ahead + sub);
            hipError_t e = rtm_hip::launch_batch_step(key_half(key), g.family, g.tile_x, key_snap16(key), b->d_jobs,
                                                      b->d_targs, g.d_tile_job, g.ntiles, sub, g.lds_bytes,
                                                      stream_of(i), lean);
            if (e != hipSuccess) return b->hip_fail(e, "batched step launch");
        }
    }
    if (nstreams > 1) {
        hipError_t e = hipSuccess;
        for (int k = 1; k < nstreams && e == hipSuccess; ++k) {
            e = hipEventRecord(b->join_ev[k - 1], b->group_streams[k - 1]);
            if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join_ev[k - 1], 0);
        }
        if (e != hipSuccess) return b->hip_fail(e, "group stream join");
    }
    hipError_t e = rtm_hip::launch_batch_advance(b->d_jobs, (int)b->active.size(), ticks, b->stream);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "clock launch");
}

// Kernel family, tile width and wavefronts per CU per group, picked by
// timing dry steps (no clock advance, no image or snapshot update, p^{n+1}
// updated in place in a scratch field, so the live state is untouched) on the
// live jobs; a candidate replaces the best when it is 2 % faster.
int autotune_merged(rtm_hip_batch* b);

int autotune(rtm_hip_batch* b) {
    if (b->cfg.kernel_family != 0) {
        for (auto& [key, g] : b->groups) {
            if (g.tuned) continue;
            g.family = static_cast<rtm_hip::Family>(b->cfg.kernel_family);
            g.tile_x = b->cfg.tile_x ? b->cfg.tile_x : (g.family == rtm_hip::Family::kStream ? 16 : 8);
            g.tuned = true;
            b->dirty = true;
        }
        if (b->merged_on && !b->merged.tuned) {
            const std::vector<int> w = rtm_hip::merged_tile_widths();
            b->merged.tile_x = std::count(w.begin(), w.end(), b->cfg.tile_x) ? b->cfg.tile_x : 4;
            b->merged.tuned = true;
            b->dirty = true;
        }
        return RTM_HIP_OK;
    }
    if (b->merged_on) return autotune_merged(b);
    bool any = false;
    for (auto& [key, g] : b->groups) any |= g.ntiles > 0 && !g.tuned;
    if (!any) return RTM_HIP_OK;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    int rc = e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "autotune events");
    // A group whose jobs run on the per-shot step kernels (run_solo,
    // run_concurrent) is timed on those, launched as those paths launch
    // them: every job of the group takes dry steps of its next eight steps
    // (skipping a 16-bit job's bookkeeping tick) per repetition, so the
    // timing sees the loop's mix of lean, storing and imaging steps.
    const bool per_shot = per_shot_kernels(b);
    auto time_per_shot = [&](int key, const Group& g, int rounds, int reps, float cutoff) -> float {
        struct Launch {
            rtm_hip::StepArgs a;
            rtm_hip::Dir dir;
            int half;
            rtm_hip::StepGeometry geom;
        };
        std::vector<Launch> launches;
        for (const Job* j : b->active) {
            if (group_key(*j->model) != key) continue;
            const Model& m = *j->model;
            const bool s16 = key_snap16(key);
            const int done = j->desc.steps - j->steps_left;
            for (int step = done; step < std::min(done + 8, j->desc.steps); ++step) {
                if (s16 && step == m.nt) continue;
                Launch l;
                l.a = s16 ? rtm_hip::job_args<true>(j->desc, step, 1) : rtm_hip::job_args<false>(j->desc, step, 1);
                l.dir = step < m.nt ? rtm_hip::Dir::Forward : rtm_hip::Dir::Backward;
                l.half = m.half;
                l.geom = rtm_hip::step_geometry(g.family, m.half, g.tile_x, m.nxe, m.nze, m.ldz);
                l.geom.blocks_per_cu = g.blocks_per_cu;
                l.geom.lds_bytes = g.lds_bytes;
                launches.push_back(l);
            }
        }
        // a per-shot launch is short: rounds of at least 48 launches, as the
        // plain engine's autotuner times them, resolve a 2 % difference
        const int n = std::max<int>(reps, (48 + (int)launches.size() - 1) / std::max<int>((int)launches.size(), 1));
        float ms = 1e30f;
        for (int round = 0; round < rounds && !rc; ++round) {
            for (int rep = 0; rep < (round == 0 ? n + 2 : n) && !rc; ++rep) {
                if (rep == (round == 0 ? 2 : 0)) (void)hipEventRecord(t0, b->stream);
                for (const Launch& l : launches) {
                    hipError_t le = rtm_hip::launch_step(l.dir, l.half, l.geom, l.a, b->stream);
                    if (le != hipSuccess) {
                        rc = b->hip_fail(le, "autotune launch");
                        break;
                    }
                }
            }
            if (rc) break;
            (void)hipEventRecord(t1, b->stream);
            hipError_t se = hipEventSynchronize(t1);
            if (se != hipSuccess) {
                rc = b->hip_fail(se, "autotune");
                break;
            }
            float round_ms = 0.0f;
            (void)hipEventElapsedTime(&round_ms, t0, t1);
            ms = std::min(ms, round_ms);
            if (ms > cutoff) break;
        }
        return ms;
    };
    // 2 warm dry steps, then the best of `rounds` rounds of `reps` timed dry steps
    // A candidate whose best round so far is above `cutoff` (well behind the
    // best candidate) cannot win: its remaining rounds are skipped.
    auto time_group = [&](int key, const Group& g, int rounds, int reps, float cutoff) -> float {
        if (per_shot) return time_per_shot(key, g, rounds, reps, cutoff);
        float ms = 1e30f;
        // dry tick arguments for tick 0 of the current clocks (the next run
        // rewrites the table)
        if (hipError_t ea = rtm_hip::launch_tick_args(b->d_jobs, (int)b->active.size(), b->d_targs, 1, 1,
                                                      b->stream);
            ea != hipSuccess)
            rc = b->hip_fail(ea, "tick arguments");
        for (int round = 0; round < rounds && !rc; ++round) {
            for (int rep = 0; rep < (round == 0 ? reps + 2 : reps) && !rc; ++rep) {
                if (rep == (round == 0 ? 2 : 0)) (void)hipEventRecord(t0, b->stream);
                hipError_t le = rtm_hip::launch_batch_step(key_half(key), g.family, g.tile_x, key_snap16(key),
                                                           b->d_jobs, b->d_targs, g.d_tile_job, g.ntiles, 0,
                                                           g.lds_bytes, b->stream);
                if (le != hipSuccess) rc = b->hip_fail(le, "autotune launch");
            }
            if (rc) break;
            (void)hipEventRecord(t1, b->stream);
            hipError_t se = hipEventSynchronize(t1);
            if (se != hipSuccess) {
                rc = b->hip_fail(se, "autotune");
                break;
            }
            float round_ms = 0.0f;
            (void)hipEventElapsedTime(&round_ms, t0, t1);
            ms = std::min(ms, round_ms);
            if (ms > cutoff) break;
        }
        return ms;
    };
    for (auto& [key, g] : b->groups) {
        if (rc || g.ntiles == 0 || g.tuned) continue;
        // The first search of a group tries every mapping and occupancy cap
        // (3 rounds of 8 steps); a new load class of a group already tuned
        // holds jobs mid-stream, so it only re-picks the DPP march width
        // (2 rounds of 4 steps), which is where the classes differ.
        const bool full = g.by_load.empty();
        const int rounds = full ? 3 : 2, reps = full ? 8 : 4;
        // stage 1: mapping and tile width at full occupancy
        float best = 1e30f;
        Tuning pick = g;
        // (per-shot kernels: the DPP march widths only, with no occupancy
        // stage, which is where the plain engine's search always ends on
        // the grids a job runs alone or beside smaller ones)
        std::vector<rtm_hip::Family> families{rtm_hip::Family::kMarchDpp};
        if (full && !per_shot)
            families = {rtm_hip::Family::kMarchDpp, rtm_hip::Family::kStream, rtm_hip::Family::kMarch};
        for (auto fam : families) {
            // batch kernel: widest first, since with many jobs per launch there
            // are waves to spare and wide tiles re-read less halo (a narrower
            // one must win by 2 %); per-shot kernels: the plain engine's order
            // of preference (4 columns first on the DPP march)
            std::vector<int> widths = rtm_hip::tile_widths(fam);
            if (!per_shot)
                std::sort(widths.begin(), widths.end(), [](int x, int y) {
                    return (x == 16) != (y == 16) ? y == 16 : x > y;  // 16 last: it spills registers
                });
            for (int tx : widths) {
                if (rc) break;
                g.tile_x = tx;
                g.family = fam;
                g.lds_bytes = 0;
                g.blocks_per_cu = 0;
                if ((rc = rebuild(b))) break;
                const float ms = time_group(key, g, rounds, reps, 1.25f * best);
                if (rc) break;
                if (ms < 0.98f * best) {
                    best = ms;
                    pick = g;
                }
            }
        }
        // stage 2: wavefronts per CU for the winner
        g.tile_x = pick.tile_x;
        g.family = pick.family;
        if (full && !per_shot && !rc && (rc = rebuild(b)) == RTM_HIP_OK) {
            for (int cap : {4, 2}) {
                g.blocks_per_cu = cap;
                g.lds_bytes = rtm_hip::occupancy_lds_bytes(b->device, cap);
                if (g.lds_bytes == 0) continue;
                const float ms = time_group(key, g, rounds, reps, 1.25f * best);
                if (rc) break;
                if (ms < 0.98f * best) {
                    best = ms;
                    pick = g;
                }
            }
        }
        static_cast<Tuning&>(g) = pick;
        g.by_load[g.load_class] = pick;
        g.tuned = true;
        b->stats.tunes += 1;
        b->dirty = true;
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    return rc;
}

// The merged launch's tile width (DPP march, merged_tile_widths()), timed on
// dry ticks of the live jobs as autotune times a group's candidates: widest
// first, a narrower width must win by 2 %; the choice is kept per load class.
int autotune_merged(rtm_hip_batch* b) {
    Group& g = b->merged;
    if (g.tuned || g.ntiles == 0) return RTM_HIP_OK;
    hipEvent_t t0 = nullptr, t1 = nullptr;
    hipError_t e = hipEventCreate(&t0);
    if (e == hipSuccess) e = hipEventCreate(&t1);
    int rc = e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "autotune events");
    const int rounds = g.by_load.empty() ? 3 : 2, reps = g.by_load.empty() ? 8 : 4;
    std::vector<int> widths = rtm_hip::merged_tile_widths();
    std::sort(widths.begin(), widths.end(), [](int x, int y) { return x > y; });
    float best = 1e30f;
    int pick = g.tile_x;
    for (int tx : widths) {
        if (rc) break;
        g.tile_x = tx;
        if ((rc = rebuild(b))) break;
        if (hipError_t ea = rtm_hip::launch_tick_args(b->d_jobs, (int)b->active.size(), b->d_targs, 1, 1, b->stream);
            ea != hipSuccess) {
            rc = b->hip_fail(ea, "tick arguments");
            break;
        }
        float ms = 1e30f;
        for (int round = 0; round < rounds && !rc; ++round) {
            for (int rep = 0; rep < (round == 0 ? reps + 2 : reps) && !rc; ++rep) {
                if (rep == (round == 0 ? 2 : 0)) (void)hipEventRecord(t0, b->stream);
                hipError_t le = rtm_hip::launch_batch_merged(g.tile_x, b->d_jobs, b->d_targs, g.d_tile_job, g.ntiles,
                                                             0, b->stream);
                if (le != hipSuccess) rc = b->hip_fail(le, "autotune launch");
            }
            if (rc) break;
            (void)hipEventRecord(t1, b->stream);
            if (hipError_t se = hipEventSynchronize(t1); se != hipSuccess) {
                rc = b->hip_fail(se, "autotune");
                break;
            }
            float round_ms = 0.0f;
            (void)hipEventElapsedTime(&round_ms, t0, t1);
            ms = std::min(ms, round_ms);
        }
        if (!rc && ms < 0.98f * best) {
            best = ms;
            pick = tx;
        }
    }
    if (t0) (void)hipEventDestroy(t0);
    if (t1) (void)hipEventDestroy(t1);
    if (rc) return rc;
    g.tile_x = pick;
    g.family = rtm_hip::Family::kMarchDpp;
    g.by_load[g.load_class] = g;
    g.tuned = true;
    b->stats.tunes += 1;
    b->dirty = true;
    return RTM_HIP_OK;
}

int admit_from_queue(rtm_hip_batch* b) {
    // jobs admitted ahead join first (they were ahead in the queue)
    size_t joined = 0;
    for (; joined < b->staged.size(); ++joined) {
        if (b->cfg.max_active > 0 && (int)b->active.size() >= b->cfg.max_active) break;
        Job* j = b->staged[joined];
        j->state = RTM_HIP_JOB_RUNNING;
        j->t_admit = now_s();
        b->active.push_back(j);
        b->dirty = true;
    }
    b->staged.erase(b->staged.begin(), b->staged.begin() + (ptrdiff_t)joined);
    if (!b->staged.empty()) return RTM_HIP_OK;
    while (!b->queue.empty()) {
        Job* j = b->queue.front();
        if (b->cfg.max_active > 0 && (int)b->active.size() >= b->cfg.max_active) break;
        const size_t need = admission_bytes(*j);
        if (b->used + need > b->budget) {
            if (b->active.empty())
                return b->fail(RTM_HIP_ERR_MEMORY,
                               "job %lld needs %.3f GiB with no job active; %.3f GiB of the %.3f GiB budget "
                               "are held by resident models", j->id, need / kGiB, b->used / kGiB, b->budget / kGiB);
            break;
        }
        if (int rc = admit_job(b, *j)) {
            job_free(b, *j);
            // a reserved arena can be too fragmented for the job while the
            // budget has room: it waits for running jobs to free memory
            if (rc == RTM_HIP_ERR_MEMORY && b->arena.active() && !b->active.empty()) break;
            b->queue.pop_front();
            j->state = RTM_HIP_JOB_FAILED;
            j->model->refs -= 1;
            model_maybe_free(b, j->model);
            return rc;
        }
        b->queue.pop_front();
        b->used += j->bytes;
        j->state = RTM_HIP_JOB_RUNNING;
        j->t_admit = now_s();
        b->active.push_back(j);
        b->dirty = true;
    }
    return RTM_HIP_OK;
}

// Admission ahead, called with the ticks up to the next job's finish just
// enqueued: queue-head jobs that will join the active set at that change
// (within the active-job cap, counting the finishing jobs out) and fit in the
// memory free now are admitted while the device runs those ticks, instead of
// on an idle device after the change.  Their uploads and clears are
// synchronous copies and fills on the null stream, which the batch's
// non-blocking streams do not wait for, into memory no queued tick touches
// (the finishing jobs keep theirs until they retire).  A job that does not
// fit, or whose admission fails, stays queued for the ordinary admission,
// which reports any error.
int stage_ahead(rtm_hip_batch* b) {
    int finishing = 0;
    for (const Job* j : b->active) finishing += j->steps_left == 0;
    if (finishing == 0) return RTM_HIP_OK;
    const double t0 = now_s();
    while (!b->queue.empty()) {
        const int after = (int)b->active.size() - finishing + (int)b->staged.size();
        if (b->cfg.max_active > 0 && after >= b->cfg.max_active) break;
        Job* j = b->queue.front();
        if (b->used + admission_bytes(*j) > b->budget) break;
        if (admit_job(b, *j) != RTM_HIP_OK) {
            job_free(b, *j);
            b->err.clear();
            break;
        }
        b->queue.pop_front();
        b->used += j->bytes;
        b->staged.push_back(j);
        b->stats.staged_admissions += 1;
    }
    b->stats.admit_seconds += now_s() - t0;
    return RTM_HIP_OK;
}

int retire_finished(rtm_hip_batch* b) {
    bool any = false;
    for (Job* j : b->active) any |= j->steps_left == 0;
    if (!any) return RTM_HIP_OK;
    std::vector<Job*> keep;
    for (Job* j : b->active) {
        if (j->steps_left > 0) {
            keep.push_back(j);
            continue;
        }
        const Model& m = *j->model;
        hipError_t e = hipSuccess;
        if (j->stack_ix0 >= 0 && b->d_stack_i)
            e = rtm_hip::launch_stack_add(j->desc.image, j->desc.illum, m.grid.nx, m.grid.nz, b->d_stack_i,
                                          b->d_stack_s, b->stack_nz, j->stack_ix0, j->stack_iz0, b->stream);
        j->image.resize(m.n_int);
        j->illum.resize(m.n_int);
        if (e == hipSuccess)
            e = hipMemcpyAsync(j->image.data(), j->desc.image, m.n_int * sizeof(float), hipMemcpyDeviceToHost,
                               b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(j->illum.data(), j->desc.illum, m.n_int * sizeof(float), hipMemcpyDeviceToHost,
                               b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) return b->hip_fail(e, "image download");
        job_free(b, *j);
        b->used -= j->bytes;
        j->state = RTM_HIP_JOB_DONE;
        j->t_done = now_s();
        j->model->refs -= 1;
        model_maybe_free(b, j->model);
        b->stats.jobs_done += 1;
        b->stats.cell_steps += 2.0 * m.nt * (double)m.nxe * (double)m.nze;
        b->stats.busy_latency_s += j->t_done - j->t_submit;
    }
    b->active.swap(keep);
    b->dirty = true;
    return RTM_HIP_OK;
}

// The captured graph of a run of `ticks` ticks starting `ahead` ticks after
// the host's clocks: a graph of the same launch pattern if one was captured
// since the job set last changed, else a new capture (the tick arguments are
// computed on the device from the job clocks, so a graph serves every run of
// its pattern).  At most kMaxGraphs patterns are kept.
constexpr size_t kMaxGraphs = 8;
int capture_ticks_now(rtm_hip_batch* b, int ticks, int ahead, rtm_hip_batch::Captured& c);

int graph_for(rtm_hip_batch* b, int ticks, int ahead, hipGraphExec_t* exec) {
    std::string pattern = run_pattern(b, ticks, ahead);
    for (const auto& c : b->graphs)
        if (c.pattern == pattern) {
            *exec = c.exec;
            return RTM_HIP_OK;
        }
    if (b->graphs.size() >= kMaxGraphs) {
        // the kept graphs may still run on the stream
        if (hipError_t e = hipStreamSynchronize(b->stream); e != hipSuccess)
            return b->hip_fail(e, "batch synchronisation");
        b->drop_graph();
    }
    rtm_hip_batch::Captured c;
    c.pattern = std::move(pattern);
    const double t0 = now_s();
    const int rc = capture_ticks_now(b, ticks, ahead, c);
    b->stats.capture_seconds += now_s() - t0;
    if (rc) return rc;
    ++b->stats.graphs_captured;
    b->graphs.push_back(c);
    *exec = c.exec;
    return RTM_HIP_OK;
}

int capture_ticks_now(rtm_hip_batch* b, int ticks, int ahead, rtm_hip_batch::Captured& c) {
    hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return b->hip_fail(e, "graph capture");
    const int rc = enqueue_ticks(b, ticks, ahead);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(b->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (e != hipSuccess) return b->hip_fail(e, "graph capture");
    c.graph = g;
    e = hipGraphInstantiate(&c.exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return b->hip_fail(e, "graph instantiation");
    }
    return RTM_HIP_OK;
}

// Ticks between job-set changes replay captured graphs: kGraphTicks ticks per
// launch while that many remain (one host call for many small ticks), then
// single ticks.
constexpr int kGraphTicks = rtm_hip::kRunTicks;

// A job alone on the device runs on the per-shot step kernels, with its
// step arguments built on the host as the per-shot engine builds them: the
// batch kernel's own cost, reading the wave's job and tick arguments from
// memory, is ~1.5 us a tick, which only many jobs per launch hide.  The same
// tile function runs either way, so the results are the same bits.  The
// bookkeeping tick of a 16-bit job still runs on the batch kernel.
int run_solo(rtm_hip_batch* b, int n) {
    Job& j = *b->active.front();
    const Model& m = *j.model;
    const int key = group_key(m);
    const Group& g = b->groups[key];
    rtm_hip::StepGeometry geom = rtm_hip::step_geometry(g.family, m.half, g.tile_x, m.nxe, m.nze, m.ldz);
    geom.blocks_per_cu = g.blocks_per_cu;
    geom.lds_bytes = g.lds_bytes;
    const bool s16 = key_snap16(key);
    const int done0 = j.desc.steps - j.steps_left;
    int synced = 0;                                  // ticks already added to the device clock
    for (int i = 0; i < n; ++i) {
        const int done = done0 + i;
        hipError_t e = hipSuccess;
        if (s16 && done == m.nt) {
            e = rtm_hip::launch_batch_advance(b->d_jobs, 1, i - synced, b->stream);
            synced = i;
            if (e == hipSuccess) e = rtm_hip::launch_tick_args(b->d_jobs, 1, b->d_targs, 1, 0, b->stream);
            if (e == hipSuccess)
                e = rtm_hip::launch_batch_step(m.half, g.family, g.tile_x, true, b->d_jobs, b->d_targs,
                                               g.d_tile_job, g.ntiles, 0, g.lds_bytes, b->stream);
        } else {
            const rtm_hip::StepArgs a = s16 ? rtm_hip::job_args<true>(j.desc, done, 0)
                                            : rtm_hip::job_args<false>(j.desc, done, 0);
            e = rtm_hip::launch_step(done < m.nt ? rtm_hip::Dir::Forward : rtm_hip::Dir::Backward, m.half, geom, a,
                                     b->stream);
        }
        if (e != hipSuccess) return b->hip_fail(e, "step launch");
    }
    hipError_t e = rtm_hip::launch_batch_advance(b->d_jobs, 1, n - synced, b->stream);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "clock launch");
}

// Two or three jobs, the largest of which fills the chip on its own: each
// job runs on the per-shot step kernels on a stream of its own, with its
// step arguments built on the host, so the smaller jobs' launches fill the
// large one's launch gaps instead of the large job paying the batch kernel's
// argument reads on every tick.  (Two chip-filling jobs side by side run
// slower than one after the other, since their fields no longer fit the
// caches; keeping them apart is the admission's business.)  The same tile
// function runs either way, so the results are the same bits.
constexpr double kConcurrentCells = 1.0e6;

bool run_concurrently(const rtm_hip_batch* b) {
    if (b->merged_on || b->active.size() < 2 || b->active.size() > (size_t)kGroupStreams) return false;
    double largest = 0.0;
    for (const Job* j : b->active) largest = std::max(largest, (double)j->model->nxe * (double)j->model->nze);
    return largest >= kConcurrentCells;
}

int run_concurrent(rtm_hip_batch* b, int n) {
    const size_t nj = b->active.size();
    auto stream_of = [&](size_t k) { return k == 0 ? b->stream : b->group_streams[k - 1]; };
    struct PerJob {
        const Job* job;
        rtm_hip::StepGeometry geom;
        int tile_x;
        bool s16;
        int done0;
    };
    std::vector<PerJob> per;
    for (const Job* j : b->active) {
        const Model& m = *j->model;
        const int key = group_key(m);
        const Group& g = b->groups[key];
        rtm_hip::StepGeometry geom = rtm_hip::step_geometry(g.family, m.half, g.tile_x, m.nxe, m.nze, m.ldz);
        geom.blocks_per_cu = g.blocks_per_cu;
        geom.lds_bytes = g.lds_bytes;
        per.push_back({j, geom, g.tile_x, key_snap16(key), j->desc.steps - j->steps_left});
    }
    hipError_t e = hipEventRecord(b->fork_ev, b->stream);
    for (size_t k = 1; k < nj && e == hipSuccess; ++k) e = hipStreamWaitEvent(b->group_streams[k - 1], b->fork_ev, 0);
    if (e != hipSuccess) return b->hip_fail(e, "job stream fork");
    // tick by tick, so the streams advance together
    for (int i = 0; i < n; ++i)
        for (size_t k = 0; k < nj; ++k) {
            const PerJob& p = per[k];
            const Model& m = *p.job->model;
            const int done = p.done0 + i;
            if (p.s16 && done == m.nt) {
                e = rtm_hip::launch_bookkeeping(b->d_jobs, (int)k, m.half, p.tile_x, p.job->desc.ntiles, stream_of(k));
            } else {
                const rtm_hip::StepArgs a = p.s16 ? rtm_hip::job_args<true>(p.job->desc, done, 0)
                                                  : rtm_hip::job_args<false>(p.job->desc, done, 0);
                const rtm_hip::Dir dir = done < m.nt ? rtm_hip::Dir::Forward : rtm_hip::Dir::Backward;
                e = rtm_hip::launch_step(dir, m.half, p.geom, a, stream_of(k));
                b->stats.step_launches += 1;
                b->stats.lean_launches += rtm_hip::step_is_lean(dir, a, p.s16) ? 1 : 0;
            }
            if (e != hipSuccess) return b->hip_fail(e, "step launch");
        }
    for (size_t k = 1; k < nj && e == hipSuccess; ++k) {
        e = hipEventRecord(b->join_ev[k - 1], b->group_streams[k - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(b->stream, b->join_ev[k - 1], 0);
    }
    if (e != hipSuccess) return b->hip_fail(e, "job stream join");
    e = rtm_hip::launch_batch_advance(b->d_jobs, (int)nj, n, b->stream);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "clock launch");
}

// Step launches of a run of ticks, and how many ran the lean build (a
// merged launch counts as lean when every job's build is).
void count_launches(rtm_hip_batch* b, int ticks, int ahead) {
    if (b->merged_on) {
        b->stats.step_launches += ticks;
        for (int sub = 0; sub < ticks; ++sub) b->stats.lean_launches += merged_tick_lean(b, ahead + sub) ? 1 : 0;
        return;
    }
    const std::string p = run_pattern(b, ticks, ahead);
    const size_t body = p.find(':') + 1;
    b->stats.step_launches += (long long)(p.size() - body);
    b->stats.lean_launches += (long long)std::count(p.begin() + (ptrdiff_t)body, p.end(), '1');
}

// Whether the host's launches would set the pace of the active set's ticks:
// a tick of fewer extended cells than about 3M per launch group (~10 us of
// device time at the step kernels' rate) is over before the host has issued
// the next tick's launches, and a captured run issues them all at once.
// Larger ticks run no slower from direct launches, and some faster: 16
// crops 600 cells wide (10M cells a tick) take 157 ms for 4000 ticks from
// direct launches against 165 ms from graph replay.
bool launch_bound(const rtm_hip_batch* b) {
    double cells = 0.0;
    for (const Job* j : b->active) cells += (double)j->model->nxe * (double)j->model->nze;
    int groups = 0;
    for (const auto& [key, g] : b->groups) groups += g.ntiles > 0;
    return cells < 3.0e6 * std::max(b->merged_on ? 1 : groups, 1);
}

int run_ticks(rtm_hip_batch* b, int n) {
    if (b->active.size() == 1) {
        if (int rc = run_solo(b, n)) return rc;
    } else if (run_concurrently(b)) {
        if (int rc = run_concurrent(b, n)) return rc;
        b->stats.concurrent_ticks += n;
    } else if (b->cfg.use_graph && launch_bound(b)) {
        for (int ahead = 0; ahead < n;) {
            const int ticks = n - ahead >= kGraphTicks ? kGraphTicks : 1;
            hipGraphExec_t exec = nullptr;
            if (int rc = graph_for(b, ticks, ahead, &exec)) return rc;
            hipError_t e = hipGraphLaunch(exec, b->stream);
            if (e != hipSuccess) return b->hip_fail(e, "graph launch");
            count_launches(b, ticks, ahead);
            ahead += ticks;
        }
    } else {
        for (int ahead = 0; ahead < n; ahead += kGraphTicks) {
            const int ticks = std::min(n - ahead, kGraphTicks);
            if (int rc = enqueue_ticks(b, ticks, ahead)) return rc;
            count_launches(b, ticks, ahead);
        }
    }
    for (Job* j : b->active) j->steps_left -= n;
    b->stats.ticks += n;
    return RTM_HIP_OK;
}

int check(const rtm_hip_batch* b) {
    if (!b) {
        g_batch_null_error = "null rtm_hip_batch handle";
        return RTM_HIP_ERR_ARG;
    }
    return RTM_HIP_OK;
}

int enqueue_job(rtm_hip_batch* b, std::unique_ptr<Job> j, long long* id) {
    if (admission_bytes(*j) > b->budget)
        return b->fail(RTM_HIP_ERR_MEMORY, "job needs %.3f GiB, more than the %.3f GiB budget",
                       admission_bytes(*j) / kGiB, b->budget / kGiB);
    j->id = b->next_job++;
    j->t_submit = now_s();
    j->model->refs += 1;
    b->queue.push_back(j.get());
    if (id) *id = j->id;
    b->jobs[j->id] = std::move(j);
    b->stats.jobs_submitted += 1;
    return RTM_HIP_OK;
}

}  // namespace

extern "C" {

int rtm_hip_batch_create(int device, const rtm_hip_batch_config* cfg, rtm_hip_batch** out) {
    if (!out) {
        g_batch_null_error = "null output pointer";
        return RTM_HIP_ERR_ARG;
    }
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        g_batch_null_error = "no HIP device available";
        return RTM_HIP_ERR_DEVICE;
    }
    if (device < 0 || device >= n || hipSetDevice(device) != hipSuccess) {
        g_batch_null_error = "device index out of range";
        return RTM_HIP_ERR_ARG;
    }
    auto b = std::make_unique<rtm_hip_batch>();
    b->device = device;
    if (cfg) b->cfg = *cfg;
    if (b->cfg.max_active < 0) {
        g_batch_null_error = "max_active must be >= 0";
        return RTM_HIP_ERR_ARG;
    }
    const int fam = b->cfg.kernel_family, tx = b->cfg.tile_x;
    const bool fam_ok = fam == 0 || fam == RTM_HIP_STEP_MARCH || fam == RTM_HIP_STEP_MARCH_DPP ||
                        fam == RTM_HIP_STEP_STREAM;
    bool tx_ok = tx == 0;
    if (fam != 0)
        for (int t : rtm_hip::tile_widths(static_cast<rtm_hip::Family>(fam))) tx_ok |= t == tx;
    if (!fam_ok || !tx_ok) {
        g_batch_null_error = "batch kernel family must be 0 (autotuned), march, march-dpp or stream, "
                             "with a tile width that family has";
        return RTM_HIP_ERR_ARG;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        g_batch_null_error = "cannot query device memory";
        return RTM_HIP_ERR_DEVICE;
    }
    b->budget = b->cfg.memory_budget ? b->cfg.memory_budget : (size_t)(0.9 * (double)free_b);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        g_batch_null_error = "stream creation failed";
        return RTM_HIP_ERR_DEVICE;
    }
    bool ok = hipEventCreateWithFlags(&b->fork_ev, hipEventDisableTiming) == hipSuccess;
    for (auto& ev : b->chunk_ev) ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < kGroupStreams - 1 && ok; ++k)
        ok = hipStreamCreateWithFlags(&b->group_streams[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->join_ev[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        g_batch_null_error = "stream creation failed";
        rtm_hip_batch_destroy(b.release());
        return RTM_HIP_ERR_DEVICE;
    }
    *out = b.release();
    return RTM_HIP_OK;
}

void rtm_hip_batch_destroy(rtm_hip_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    b->drop_graph();
    for (Job* j : b->active) job_free(b, *j);
    for (Job* j : b->staged) job_free(b, *j);
    for (auto& [id, j] : b->jobs) model_release(b, *j->model);
    for (auto& [id, m] : b->models) model_release(b, *m);
    for (auto& [key, g] : b->groups)
        if (g.d_tile_job) (void)hipFree(g.d_tile_job);
    if (b->merged.d_tile_job) (void)hipFree(b->merged.d_tile_job);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_targs) (void)hipFree(b->d_targs);
    if (b->d_zero_ptr) (void)hipFree(b->d_zero_ptr);
    if (b->d_stack_i) (void)hipFree(b->d_stack_i);
    if (b->d_stack_s) (void)hipFree(b->d_stack_s);
    if (b->arena.active()) (void)hipFree(b->arena.base());
    for (int k = 0; k < kGroupStreams - 1; ++k) {
        if (b->group_streams[k]) {
            (void)hipStreamSynchronize(b->group_streams[k]);
            (void)hipStreamDestroy(b->group_streams[k]);
        }
        if (b->join_ev[k]) (void)hipEventDestroy(b->join_ev[k]);
    }
    if (b->fork_ev) (void)hipEventDestroy(b->fork_ev);
    for (hipEvent_t ev : b->chunk_ev)
        if (ev) (void)hipEventDestroy(ev);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int rtm_hip_batch_reserve(rtm_hip_batch* b, size_t bytes) {
    if (int rc = check(b)) return rc;
    if (!b->active.empty() || !b->queue.empty() || !b->staged.empty())
        return b->fail(RTM_HIP_ERR_ARG, "reserve memory before submitting jobs");
    if (b->arena.active()) return b->fail(RTM_HIP_ERR_ARG, "memory is already reserved");
    if (bytes == 0) bytes = b->budget > b->used ? b->budget - b->used : 0;
    bytes &= ~size_t(255);
    if (bytes == 0) return b->fail(RTM_HIP_ERR_MEMORY, "no budget left to reserve");
    (void)hipSetDevice(b->device);
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) return b->hip_fail(e, "memory reservation");
    b->arena.reset(static_cast<char*>(p), bytes);
    b->budget = b->used + bytes;
    return RTM_HIP_OK;
}

const char* rtm_hip_batch_last_error(const rtm_hip_batch* b) {
    return b ? b->err.c_str() : g_batch_null_error.c_str();
}

int rtm_hip_batch_add_model(rtm_hip_batch* b, const rtm_hip_grid* grid, const float* v, const rtm_hip_params* params,
                            int nt, float dt, int* model_id) {
    if (int rc = check(b)) return rc;
    auto m = std::make_shared<Model>();
    if (int rc = model_init(b, *m, grid, v, params, nt, dt)) return rc;
    if (b->used + m->bytes > b->budget)
        return b->fail(RTM_HIP_ERR_MEMORY, "model needs %.3f GiB; %.3f GiB of the %.3f GiB budget are in use",
                       m->bytes / kGiB, b->used / kGiB, b->budget / kGiB);
    (void)hipSetDevice(b->device);
    if (int rc = model_upload(b, *m)) return rc;
    m->registered = true;
    m->id = b->next_model++;
    b->models[m->id] = m;
    if (model_id) *model_id = m->id;
    return RTM_HIP_OK;
}

int rtm_hip_batch_remove_model(rtm_hip_batch* b, int model_id) {
    if (int rc = check(b)) return rc;
    auto it = b->models.find(model_id);
    if (it == b->models.end()) return b->fail(RTM_HIP_ERR_ARG, "unknown model %d", model_id);
    it->second->removing = true;
    model_maybe_free(b, it->second);
    b->models.erase(it);
    return RTM_HIP_OK;
}

int rtm_hip_batch_job_bytes(rtm_hip_batch* b, const rtm_hip_job* job, size_t* bytes) {
    if (int rc = check(b)) return rc;
    if (!job || !bytes) return b->fail(RTM_HIP_ERR_ARG, "null job");
    auto m = std::make_shared<Model>();
    if (int rc = model_init(b, *m, &job->grid, job->v, &job->params, job->nt, job->dt)) return rc;
    Job j;
    if (int rc = job_init(b, j, m, &job->shot, -1, -1)) return rc;
    *bytes = admission_bytes(j);
    return RTM_HIP_OK;
}

int rtm_hip_batch_submit(rtm_hip_batch* b, const rtm_hip_job* job, long long* id) {
    if (int rc = check(b)) return rc;
    if (!job) return b->fail(RTM_HIP_ERR_ARG, "null job");
    auto m = std::make_shared<Model>();
    if (int rc = model_init(b, *m, &job->grid, job->v, &job->params, job->nt, job->dt)) return rc;
    auto j = std::make_unique<Job>();
    if (int rc = job_init(b, *j, m, &job->shot, job->stack_ix0, job->stack_iz0)) return rc;
    return enqueue_job(b, std::move(j), id);
}

int rtm_hip_batch_submit_shot(rtm_hip_batch* b, int model_id, const rtm_hip_shot* shot, int stack_ix0,
                              int stack_iz0, long long* id) {
    if (int rc = check(b)) return rc;
    auto it = b->models.find(model_id);
    if (it == b->models.end()) return b->fail(RTM_HIP_ERR_ARG, "unknown model %d", model_id);
    auto j = std::make_unique<Job>();
    if (int rc = job_init(b, *j, it->second, shot, stack_ix0, stack_iz0)) return rc;
    return enqueue_job(b, std::move(j), id);
}

int rtm_hip_batch_run(rtm_hip_batch* b, int max_ticks, int until_done) {
    if (int rc = check(b)) return rc;
    (void)hipSetDevice(b->device);
    const double t0 = now_s();
    long long ran = 0;
    int rc = RTM_HIP_OK;
    for (;;) {
        const double ta = now_s();
        rc = admit_from_queue(b);
        b->stats.admit_seconds += now_s() - ta;
        if (rc) break;
        if (b->active.empty()) break;
        if (b->dirty && (rc = rebuild(b))) break;
        // every job moves one step per tick: run until the next one finishes
        int n = b->active.front()->steps_left;
        for (Job* j : b->active) n = std::min(n, j->steps_left);
        // The autotuner times dry steps on the live jobs; a job's first two
        // steps read p^n and p^{n-1} as zero and skip those loads, so a new
        // half-width group first runs two steps with the default kernel.
        int warm = 0;
        if (b->cfg.kernel_family == 0)
            for (Job* j : b->active)
                if (b->merged_on ? !b->merged.tuned : !b->groups[group_key(*j->model)].tuned)
                    warm = std::max(warm, 2 - (j->desc.steps - j->steps_left));
        if (warm > 0) {
            n = std::min(n, warm);
        } else {
            const double tt = now_s();
            b->timing = true;
            rc = autotune(b);
            b->timing = false;
            // the searched jobs' own tables are rebuilt for the width picked
            for (const Job* j : b->active) b->dirty |= j->tile_x_built < 0;
            b->stats.tune_seconds += now_s() - tt;
            if (rc) break;
            if (b->dirty && (rc = rebuild(b))) break;
        }
        if (max_ticks > 0) n = (int)std::min<long long>(n, max_ticks - ran);
        if (n <= 0) break;
        if ((rc = run_ticks(b, n))) break;
        ran += n;
        if ((rc = stage_ahead(b))) break;
        const double tr = now_s();
        rc = retire_finished(b);
        b->stats.retire_seconds += now_s() - tr;
        if (rc) break;
        if (!until_done && (max_ticks <= 0 || ran >= max_ticks)) break;
        if (until_done && max_ticks > 0 && ran >= max_ticks) break;
    }
    // A call that leaves jobs running (until_done = 0) returns with its last
    // ticks still queued on the device and waits only for the ticks the
    // previous such call left: the device has work while the caller admits
    // new jobs between calls.  Otherwise, and on an error, it waits for all.
    hipError_t e;
    if (rc || until_done || b->active.empty()) {
        e = hipStreamSynchronize(b->stream);
        b->chunk_pending = false;
    } else {
        e = hipEventRecord(b->chunk_ev[b->chunk_cur], b->stream);
        if (e == hipSuccess && b->chunk_pending) e = hipEventSynchronize(b->chunk_ev[b->chunk_cur ^ 1]);
        b->chunk_pending = true;
        b->chunk_cur ^= 1;
    }
    b->stats.run_seconds += now_s() - t0;
    if (rc) return rc;
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "batch run");
}

int rtm_hip_batch_job_state(rtm_hip_batch* b, long long id, int* state, double* latency_s) {
    if (int rc = check(b)) return rc;
    auto it = b->jobs.find(id);
    if (it == b->jobs.end()) return b->fail(RTM_HIP_ERR_ARG, "unknown job %lld", id);
    const Job& j = *it->second;
    if (state) *state = j.state;
    if (latency_s) *latency_s = j.state == RTM_HIP_JOB_DONE ? j.t_done - j.t_submit : -1.0;
    return RTM_HIP_OK;
}

int rtm_hip_batch_result(rtm_hip_batch* b, long long id, float* image, float* illum) {
    if (int rc = check(b)) return rc;
    auto it = b->jobs.find(id);
    if (it == b->jobs.end()) return b->fail(RTM_HIP_ERR_ARG, "unknown job %lld", id);
    Job& j = *it->second;
    if (j.state != RTM_HIP_JOB_DONE) return b->fail(RTM_HIP_ERR_ARG, "job %lld has not finished", id);
    if (image) std::memcpy(image, j.image.data(), j.image.size() * sizeof(float));
    if (illum) std::memcpy(illum, j.illum.data(), j.illum.size() * sizeof(float));
    b->jobs.erase(it);
    return RTM_HIP_OK;
}

int rtm_hip_batch_set_stack(rtm_hip_batch* b, int nx, int nz) {
    if (int rc = check(b)) return rc;
    if (nx <= 0 || nz <= 0) return b->fail(RTM_HIP_ERR_ARG, "stack size must be positive");
    for (Job* j : b->queue)
        if (j->stack_ix0 >= 0) return b->fail(RTM_HIP_ERR_ARG, "queued jobs use the current stack");
    for (Job* j : b->staged)
        if (j->stack_ix0 >= 0) return b->fail(RTM_HIP_ERR_ARG, "queued jobs use the current stack");
    for (Job* j : b->active)
        if (j->stack_ix0 >= 0) return b->fail(RTM_HIP_ERR_ARG, "active jobs use the current stack");
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    if (b->d_stack_i) (void)hipFree(b->d_stack_i);
    if (b->d_stack_s) (void)hipFree(b->d_stack_s);
    b->d_stack_i = b->d_stack_s = nullptr;
    b->stack_nx = b->stack_nz = 0;
    const size_t n = (size_t)nx * nz;
    hipError_t e = hipMalloc(&b->d_stack_i, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->d_stack_s, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(b->d_stack_i, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(b->d_stack_s, 0, n * sizeof(double));
    if (e != hipSuccess) {
        if (b->d_stack_i) (void)hipFree(b->d_stack_i);
        if (b->d_stack_s) (void)hipFree(b->d_stack_s);
        b->d_stack_i = b->d_stack_s = nullptr;
        return b->hip_fail(e, "FP64 stack");
    }
    b->stack_nx = nx;
    b->stack_nz = nz;
    return RTM_HIP_OK;
}

int rtm_hip_batch_get_stack(rtm_hip_batch* b, double* image, double* illum) {
    if (int rc = check(b)) return rc;
    if (!b->d_stack_i) return b->fail(RTM_HIP_ERR_ARG, "no stack is set");
    (void)hipSetDevice(b->device);
    const size_t n = (size_t)b->stack_nx * b->stack_nz;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e == hipSuccess && image) e = hipMemcpy(image, b->d_stack_i, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && illum) e = hipMemcpy(illum, b->d_stack_s, n * sizeof(double), hipMemcpyDeviceToHost);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "stack download");
}

int rtm_hip_batch_get_stats(rtm_hip_batch* b, rtm_hip_batch_stats* out) {
    if (int rc = check(b)) return rc;
    if (!out) return RTM_HIP_OK;
    *out = b->stats;
    out->jobs_active = (int)b->active.size();
    out->jobs_queued = (int)(b->queue.size() + b->staged.size());
    out->models_resident = (int)b->models.size();
    out->bytes_in_use = b->used;
    out->bytes_budget = b->budget;
    int k = 0;
    for (auto& [key, g] : b->groups) {
        if (k >= 12) break;
        // while merged, the groups present run in the one launch, with its tiles
        const Group& run = b->merged_on ? b->merged : g;
        out->group_half[k] = key_half(key);
        out->group_snap16[k] = key_snap16(key) ? 1 : 0;
        out->group_tile_x[k] = run.tile_x;
        out->group_family[k] = (int)run.family;
        out->group_waves_per_cu[k] = run.blocks_per_cu * rtm_hip::kWaves;
        ++k;
    }
    out->ngroups = k;
    out->merged_launch = b->merged_on ? 1 : 0;
    return RTM_HIP_OK;
}

}  // extern "C"
