// batch_engine.hip -- the batched and served engine (C ABI in rtm_hip.h,
// rtm_hip_batch_*): many shot-migration jobs, each on its own grid, advance
// together one time step per tick (batch_kernels.hip).
//
// Scheduling.  Submitted jobs wait in a FIFO queue.  A job is admitted when
// its device footprint (tables, three wavefields, snapshot store, image,
// traces, injection tables) fits in what the memory budget has left and the
// active-job cap allows; admission allocates and uploads it and the launch
// tables are rebuilt.  The engine then runs ticks until the next job
// finishes (the host knows every job's clock: one step per tick), takes the
// finished job's image (and adds it into the FP64 stack when the job has a
// stack position), frees its memory and admits from the queue again: jobs
// that arrive while others are on the device start as soon as memory frees
// up, and their steps overlap with the running jobs' steps in the same
// launches (continuous batching).
//
// Launches.  Jobs are grouped by stencil half-width; each group has one
// launch per tick over the concatenation of its jobs' tiles, with the tile
// width and cross-lane method picked once per half-width by timing dry runs
// on the live jobs.  A tick (the group launches plus the clock kernel) can be
// captured once in a hipGraph and replayed until the job set changes.
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

// Host copy of one submitted job and, once admitted, its device buffers.
struct Job {
    long long id = 0;
    int state = RTM_HIP_JOB_QUEUED;
    // inputs
    rtm_hip_grid grid{};
    std::vector<float> v;
    rtm_hip_params params{};
    int nt = 0;
    float dt = 0.0f;
    float sx = 0.0f, sz = 0.0f;
    std::vector<float> rx, rz, traces;
    int trace_nt = 0;
    int stack_ix0 = -1, stack_iz0 = -1;
    // derived
    int half = 0, nxe = 0, nze = 0, ldz = 0, nsnap = 0;
    size_t n_int = 0, n_fld = 0, bytes = 0;
    rtm_hip::Coeffs coef{};
    // device
    std::vector<void*> allocs;       // freed when the job retires
    std::vector<void*> entry_allocs; // injection tables of the current tile width
    JobDesc desc{};
    int tile_x_built = 0;            // geometry the entry tables were built for
    int steps_left = 0;              // ticks until done (host mirror of the clock)
    // timing and result
    double t_submit = 0, t_admit = 0, t_done = 0;
    std::vector<float> image, illum;
};

struct Group {
    int half = 0;
    int tile_x = 4;
    bool dpp = true;
    bool tuned = false;
    int ntiles = 0;
    int* d_tile_job = nullptr;
    size_t cap = 0;
};

}  // namespace

struct rtm_hip_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    rtm_hip_batch_config cfg{};
    size_t budget = 0, used = 0;
    long long next_id = 1;
    std::map<long long, std::unique_ptr<Job>> jobs;
    std::deque<Job*> queue;
    std::vector<Job*> active;
    JobDesc* d_jobs = nullptr;
    size_t jobs_cap = 0;
    std::map<int, Group> groups;     // by half-width
    bool dirty = true;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // FP64 stack
    int stack_nx = 0, stack_nz = 0;
    double* d_stack_i = nullptr;
    double* d_stack_s = nullptr;
    // accounting
    rtm_hip_batch_stats stats{};

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
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        graph_exec = nullptr;
        graph = nullptr;
    }
};

namespace {

// Device bytes one job needs while active.
size_t job_footprint(const Job& j) {
    const size_t snap_elem = j.params.snapshot_format == RTM_HIP_SNAP_F32 ? 4 : 2;
    const size_t tiles_bound = (size_t)(j.nxe / 4 + 2) * (j.nze / 52 + 2) + 1;   // tile_x >= 4, strip >= 52
    size_t b = 0;
    b += 4 * j.n_fld * sizeof(float);                       // vdt2 + three wavefields
    b += ((size_t)j.nxe + j.ldz + j.nt) * sizeof(float);    // tapers, wavelet
    b += (size_t)j.nsnap * j.n_int * snap_elem;             // snapshot store
    b += (size_t)j.nsnap * sizeof(float);                   // snapshot scales
    b += 2 * j.n_int * sizeof(float);                       // image, illumination
    b += j.traces.size() * sizeof(float);                   // receiver data
    b += 2 * (tiles_bound * sizeof(int) + (j.rx.size() + 1) * sizeof(Entry));
    return b;
}

int derive(rtm_hip_batch* b, Job& j) {
    const int order = j.params.order;
    if (order != 2 && order != 4 && order != 6 && order != 8 && order != 10 && order != 12)
        return b->fail(RTM_HIP_ERR_ARG, "unsupported spatial order (use 2, 4, 6, 8, 10 or 12)");
    j.half = order / 2;
    if (j.params.nb < j.half + 1) return b->fail(RTM_HIP_ERR_ARG, "--nb must be at least order/2 + 1");
    if (j.grid.nx <= 0 || j.grid.nz <= 0) return b->fail(RTM_HIP_ERR_ARG, "velocity model size does not match nx*nz");
    if (j.params.store_interval < 1) return b->fail(RTM_HIP_ERR_ARG, "--store-interval must be >= 1");
    if (j.nt < 1 || !(j.dt > 0.0f)) return b->fail(RTM_HIP_ERR_ARG, "nt must be >= 1 and dt > 0");
    const int fmt = j.params.snapshot_format;
    if (fmt != RTM_HIP_SNAP_F32 && fmt != RTM_HIP_SNAP_F16 && fmt != RTM_HIP_SNAP_BF16)
        return b->fail(RTM_HIP_ERR_ARG, "unknown snapshot format");
    if (j.rx.empty()) return b->fail(RTM_HIP_ERR_ARG, "a migration job needs receivers and traces");
    if (j.trace_nt < j.nt)
        return b->fail(RTM_HIP_ERR_ARG, "shot has %d samples per trace, fewer than nt=%d", j.trace_nt, j.nt);
    j.nxe = j.grid.nx + 2 * j.params.nb;
    j.nze = j.grid.nz + 2 * j.params.nb;
    j.ldz = round_up(j.nze, rtm_hip::kTileZ);
    if ((long long)j.nxe * j.ldz >= (1LL << 31)) return b->fail(RTM_HIP_ERR_ARG, "extended grid exceeds 2^31 cells");
    j.nsnap = (j.nt - 1) / j.params.store_interval + 1;
    j.n_int = (size_t)j.grid.nx * j.grid.nz;
    j.n_fld = (size_t)j.nxe * j.ldz;
    j.bytes = job_footprint(j);
    return RTM_HIP_OK;
}

template <typename T>
int job_alloc(rtm_hip_batch* b, Job& j, T** p, size_t count, const char* what) {
    hipError_t e = hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T));
    if (e != hipSuccess) {
        *p = nullptr;
        return b->hip_fail(e, what);
    }
    j.allocs.push_back(*p);
    return RTM_HIP_OK;
}

void free_entries(Job& j) {
    for (void* p : j.entry_allocs) (void)hipFree(p);
    j.entry_allocs.clear();
    j.tile_x_built = 0;
}

void job_free(Job& j) {
    free_entries(j);
    for (void* p : j.allocs) (void)hipFree(p);
    j.allocs.clear();
}

// Nearest-grid-point cell of a position, clamped to the interior
// (src/rtm/rtm_cpu.cpp:99-123).
void job_cell(const Job& j, float x, float z, int* ixe, int* ize) {
    const int ix0 = static_cast<int>(std::lround((x - j.grid.ox) / j.grid.dx));
    const int iz0 = static_cast<int>(std::lround((z - j.grid.oz) / j.grid.dz));
    *ixe = std::clamp(ix0, 0, j.grid.nx - 1) + j.params.nb;
    *ize = std::clamp(iz0, 0, j.grid.nz - 1) + j.params.nb;
}

// Per-tile injection tables of the job for tile width tile_x (forward: the
// source; backward: every receiver, in receiver order per cell).
int build_job_entries(rtm_hip_batch* b, Job& j, int tile_x) {
    free_entries(j);
    const rtm_hip::StepGeometry g =
        rtm_hip::step_geometry(rtm_hip::Family::kMarchDpp, j.half, tile_x, j.nxe, j.nze, j.ldz);
    for (int k = 0; k < 2; ++k) {
        struct T {
            int tile, cell, code;
        };
        std::vector<T> v;
        int ixe, ize;
        if (k == 0) {
            job_cell(j, j.sx, j.sz, &ixe, &ize);
            v.push_back({rtm_hip::tile_of_cell(g, ixe, ize), ixe * j.ldz + ize, -1});
        } else {
            for (size_t ir = 0; ir < j.rx.size(); ++ir) {
                job_cell(j, j.rx[ir], j.rz[ir], &ixe, &ize);
                v.push_back({rtm_hip::tile_of_cell(g, ixe, ize), ixe * j.ldz + ize, (int)ir});
            }
        }
        std::sort(v.begin(), v.end(), [](const T& a, const T& c) {
            if (a.tile != c.tile) return a.tile < c.tile;
            if (a.cell != c.cell) return a.cell < c.cell;
            return a.code < c.code;
        });
        std::vector<int> ptr((size_t)g.ntiles + 1, 0);
        std::vector<Entry> ent(v.size());
        for (size_t i = 0; i < v.size(); ++i) {
            ent[i] = Entry{v[i].cell, v[i].code};
            ++ptr[(size_t)v[i].tile + 1];
        }
        for (int t = 0; t < g.ntiles; ++t) ptr[t + 1] += ptr[t];
        int* d_ptr = nullptr;
        Entry* d_ent = nullptr;
        if (int rc = job_alloc(b, j, &d_ptr, ptr.size(), "injection tables")) return rc;
        if (int rc = job_alloc(b, j, &d_ent, ent.size(), "injection tables")) return rc;
        j.allocs.pop_back();
        j.allocs.pop_back();
        j.entry_allocs.push_back(d_ptr);
        j.entry_allocs.push_back(d_ent);
        hipError_t e = hipMemcpy(d_ptr, ptr.data(), ptr.size() * sizeof(int), hipMemcpyHostToDevice);
        if (e == hipSuccess && !ent.empty())
            e = hipMemcpy(d_ent, ent.data(), ent.size() * sizeof(Entry), hipMemcpyHostToDevice);
        if (e != hipSuccess) return b->hip_fail(e, "injection table upload");
        j.desc.tile_ptr[k] = d_ptr;
        j.desc.entries[k] = d_ent;
    }
    j.desc.ntz = g.ntz;
    j.desc.ntx = g.ntx;
    j.desc.ntiles = g.ntiles;
    j.tile_x_built = tile_x;
    return RTM_HIP_OK;
}

// Allocate, build and upload everything an admitted job needs.
int admit_job(rtm_hip_batch* b, Job& j) {
    rtm_hip::HostTables t;
    try {
        t = rtm_hip::build_tables(j.grid.nx, j.grid.nz, j.grid.dx, j.grid.dz, j.params.nb, j.params.order,
                                  j.params.sponge_alpha, j.params.f0, j.nt, j.dt, j.v.data(), j.ldz);
    } catch (const std::exception& ex) {
        return b->fail(RTM_HIP_ERR_ARG, "%s", ex.what());
    }
    JobDesc& d = j.desc;
    d = JobDesc{};
    d.nxe = j.nxe; d.nze = j.nze; d.ldz = j.ldz; d.nb = j.params.nb; d.nx = j.grid.nx; d.nz = j.grid.nz;
    d.half = j.half; d.nt = j.nt; d.store = j.params.store_interval; d.snap_fmt = j.params.snapshot_format;
    d.coef = t.coef;
    d.trace_stride = j.trace_nt;
    d.n_int = (long long)j.n_int;
    d.bwd_base = (j.nt + 1) % 3;            // buffer of the forward phase's last field
    float *vdt2, *wx, *wz, *wav, *tr, *img, *ill;
    int rc = RTM_HIP_OK;
    if (!rc) rc = job_alloc(b, j, &vdt2, j.n_fld, "vdt2");
    if (!rc) rc = job_alloc(b, j, &wx, (size_t)j.nxe, "wx");
    if (!rc) rc = job_alloc(b, j, &wz, (size_t)j.ldz, "wz");
    if (!rc) rc = job_alloc(b, j, &wav, (size_t)j.nt, "wavelet");
    if (!rc) rc = job_alloc(b, j, &tr, j.traces.size(), "traces");
    if (!rc) rc = job_alloc(b, j, &img, j.n_int, "image");
    if (!rc) rc = job_alloc(b, j, &ill, j.n_int, "illumination");
    for (int k = 0; k < 3 && !rc; ++k) rc = job_alloc(b, j, &d.fld[k], j.n_fld, "wavefield");
    if (d.snap_fmt == RTM_HIP_SNAP_F32) {
        if (!rc) rc = job_alloc(b, j, &d.snap32, (size_t)j.nsnap * j.n_int, "snapshot store");
    } else {
        if (!rc) rc = job_alloc(b, j, &d.snap16, (size_t)j.nsnap * j.n_int, "snapshot store");
        if (!rc) rc = job_alloc(b, j, &d.snap_max, (size_t)j.nsnap, "snapshot scales");
    }
    if (rc) return rc;
    hipError_t e = hipMemcpy(vdt2, t.vdt2.data(), t.vdt2.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(wx, t.wx.data(), t.wx.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(wz, t.wz.data(), t.wz.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(wav, t.wavelet.data(), t.wavelet.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(tr, j.traces.data(), j.traces.size() * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemset(img, 0, j.n_int * sizeof(float));
    if (e == hipSuccess) e = hipMemset(ill, 0, j.n_int * sizeof(float));
    for (int k = 0; k < 3 && e == hipSuccess; ++k) e = hipMemset(d.fld[k], 0, j.n_fld * sizeof(float));
    if (e == hipSuccess && d.snap_max) e = hipMemset(d.snap_max, 0, (size_t)j.nsnap * sizeof(float));
    if (e != hipSuccess) return b->hip_fail(e, "job upload");
    d.vdt2 = vdt2; d.wx = wx; d.wz = wz; d.wavelet = wav; d.traces = tr; d.image = img; d.illum = ill;
    d.phase = 0;
    d.step = 0;
    j.steps_left = 2 * j.nt;
    j.tile_x_built = 0;
    j.traces.clear();
    j.traces.shrink_to_fit();
    j.v.clear();
    j.v.shrink_to_fit();
    return RTM_HIP_OK;
}

// The launch tables of every half-width group and the device descriptors.
int rebuild(rtm_hip_batch* b) {
    b->drop_graph();
    if (b->active.size() > b->jobs_cap) {
        if (b->d_jobs) (void)hipFree(b->d_jobs);
        b->jobs_cap = std::max<size_t>(b->active.size(), 2 * b->jobs_cap);
        hipError_t e = hipMalloc(&b->d_jobs, b->jobs_cap * sizeof(JobDesc));
        if (e != hipSuccess) {
            b->d_jobs = nullptr;
            b->jobs_cap = 0;
            return b->hip_fail(e, "job descriptors");
        }
    }
    std::map<int, std::vector<int>> by_half;
    for (size_t k = 0; k < b->active.size(); ++k) by_half[b->active[k]->half].push_back((int)k);
    for (auto it = b->groups.begin(); it != b->groups.end();) {
        if (!by_half.count(it->first)) it->second.ntiles = 0;
        ++it;
    }
    for (auto& [half, idx] : by_half) {
        Group& g = b->groups[half];
        g.half = half;
        std::vector<int> tile_job;
        for (int k : idx) {
            Job& j = *b->active[k];
            if (j.tile_x_built != g.tile_x)
                if (int rc = build_job_entries(b, j, g.tile_x)) return rc;
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
        if (e != hipSuccess) return b->hip_fail(e, "tile table upload");
    }
    // the host knows every clock: one step per tick since admission
    std::vector<JobDesc> descs;
    for (Job* j : b->active) {
        const int done = 2 * j->nt - j->steps_left;
        j->desc.phase = done / j->nt;
        j->desc.step = done % j->nt;
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

// Launches of one tick.
int enqueue_tick(rtm_hip_batch* b, int dry) {
    for (auto& [half, g] : b->groups) {
        if (g.ntiles == 0) continue;
        hipError_t e = rtm_hip::launch_batch_step(half, g.tile_x, g.dpp, b->d_jobs, g.d_tile_job, g.ntiles, dry,
                                                  b->stream);
        if (e != hipSuccess) return b->hip_fail(e, "batched step launch");
    }
    if (dry) return RTM_HIP_OK;
    hipError_t e = rtm_hip::launch_batch_advance(b->d_jobs, (int)b->active.size(), b->stream);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "clock launch");
}

// Pick tile width and cross-lane method per half-width by timing dry ticks
// (no clock advance, no image or snapshot update) on the live jobs.
int autotune(rtm_hip_batch* b) {
    bool any = false;
    for (auto& [half, g] : b->groups)
        if (g.ntiles > 0 && !g.tuned) any = true;
    if (!any) return RTM_HIP_OK;
    hipEvent_t t0, t1;
    (void)hipEventCreate(&t0);
    (void)hipEventCreate(&t1);
    for (auto& [half, g] : b->groups) {
        if (g.ntiles == 0 || g.tuned) continue;
        float best = 1e30f;
        int best_tx = g.tile_x;
        bool best_dpp = g.dpp;
        for (int tx : {4, 8, 16}) {
            for (bool dpp : {true, false}) {
                g.tile_x = tx;
                g.dpp = dpp;
                if (int rc = rebuild(b)) return rc;
                for (int rep = 0; rep < 4; ++rep) {
                    if (rep == 1) (void)hipEventRecord(t0, b->stream);
                    hipError_t e = rtm_hip::launch_batch_step(half, tx, dpp, b->d_jobs, g.d_tile_job, g.ntiles, 1,
                                                              b->stream);
                    if (e != hipSuccess) return b->hip_fail(e, "autotune launch");
                }
                (void)hipEventRecord(t1, b->stream);
                (void)hipEventSynchronize(t1);
                float ms = 0.0f;
                (void)hipEventElapsedTime(&ms, t0, t1);
                if (ms < 0.97f * best) {
                    best = ms;
                    best_tx = tx;
                    best_dpp = dpp;
                }
            }
        }
        g.tile_x = best_tx;
        g.dpp = best_dpp;
        g.tuned = true;
        b->dirty = true;
    }
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    return RTM_HIP_OK;
}

int admit_from_queue(rtm_hip_batch* b) {
    while (!b->queue.empty()) {
        Job* j = b->queue.front();
        if (b->cfg.max_active > 0 && (int)b->active.size() >= b->cfg.max_active) break;
        if (b->used + j->bytes > b->budget) {
            if (b->active.empty())
                return b->fail(RTM_HIP_ERR_MEMORY, "job %lld needs %.2f GiB, more than the %.2f GiB budget", j->id,
                               j->bytes / 1073741824.0, b->budget / 1073741824.0);
            break;
        }
        b->queue.pop_front();
        if (int rc = admit_job(b, *j)) {
            job_free(*j);
            j->state = RTM_HIP_JOB_FAILED;
            return rc;
        }
        b->used += j->bytes;
        j->state = RTM_HIP_JOB_RUNNING;
        j->t_admit = now_s();
        b->active.push_back(j);
        b->dirty = true;
    }
    return RTM_HIP_OK;
}

int retire_finished(rtm_hip_batch* b) {
    bool any = false;
    for (Job* j : b->active) any |= j->steps_left == 0;
    if (!any) return RTM_HIP_OK;
    hipError_t e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return b->hip_fail(e, "batch synchronisation");
    std::vector<Job*> keep;
    for (Job* j : b->active) {
        if (j->steps_left > 0) {
            keep.push_back(j);
            continue;
        }
        if (j->stack_ix0 >= 0 && b->d_stack_i) {
            e = rtm_hip::launch_stack_add(j->desc.image, j->desc.illum, j->grid.nx, j->grid.nz, b->d_stack_i,
                                          b->d_stack_s, b->stack_nz, j->stack_ix0, j->stack_iz0, b->stream);
            if (e != hipSuccess) return b->hip_fail(e, "FP64 stack update");
        }
        j->image.resize(j->n_int);
        j->illum.resize(j->n_int);
        e = hipMemcpyAsync(j->image.data(), j->desc.image, j->n_int * sizeof(float), hipMemcpyDeviceToHost, b->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(j->illum.data(), j->desc.illum, j->n_int * sizeof(float), hipMemcpyDeviceToHost,
                               b->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
        if (e != hipSuccess) return b->hip_fail(e, "image download");
        job_free(*j);
        b->used -= j->bytes;
        j->state = RTM_HIP_JOB_DONE;
        j->t_done = now_s();
        b->stats.jobs_done += 1;
        b->stats.cell_steps += 2.0 * j->nt * (double)j->nxe * (double)j->nze;
        b->stats.busy_latency_s += j->t_done - j->t_submit;
    }
    b->active.swap(keep);
    b->dirty = true;
    return RTM_HIP_OK;
}

int run_ticks(rtm_hip_batch* b, int n) {
    if (b->cfg.use_graph) {
        if (!b->graph_exec) {
            hipError_t e = hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return b->hip_fail(e, "graph capture");
            const int rc = enqueue_tick(b, 0);
            hipGraph_t g = nullptr;
            e = hipStreamEndCapture(b->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            if (e != hipSuccess) return b->hip_fail(e, "graph capture");
            b->graph = g;
            e = hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) return b->hip_fail(e, "graph instantiation");
        }
        for (int i = 0; i < n; ++i) {
            hipError_t e = hipGraphLaunch(b->graph_exec, b->stream);
            if (e != hipSuccess) return b->hip_fail(e, "graph launch");
        }
    } else {
        for (int i = 0; i < n; ++i)
            if (int rc = enqueue_tick(b, 0)) return rc;
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
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    b->budget = b->cfg.memory_budget ? b->cfg.memory_budget : (size_t)(0.9 * (double)free_b);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
        g_batch_null_error = "stream creation failed";
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
    for (Job* j : b->active) job_free(*j);
    for (auto& [half, g] : b->groups)
        if (g.d_tile_job) (void)hipFree(g.d_tile_job);
    if (b->d_jobs) (void)hipFree(b->d_jobs);
    if (b->d_stack_i) (void)hipFree(b->d_stack_i);
    if (b->d_stack_s) (void)hipFree(b->d_stack_s);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

const char* rtm_hip_batch_last_error(const rtm_hip_batch* b) {
    return b ? b->err.c_str() : g_batch_null_error.c_str();
}

int rtm_hip_batch_job_bytes(rtm_hip_batch* b, const rtm_hip_job* job, size_t* bytes) {
    if (int rc = check(b)) return rc;
    if (!job || !bytes) return b->fail(RTM_HIP_ERR_ARG, "null job");
    Job j;
    j.grid = job->grid;
    j.params = job->params;
    j.nt = job->nt;
    j.dt = job->dt;
    j.rx.assign(job->shot.rx, job->shot.rx + std::max(job->shot.nrec, 0));
    j.trace_nt = job->shot.nt;
    j.traces.resize((size_t)std::max(job->shot.nrec, 0) * std::max(job->shot.nt, 0));
    if (int rc = derive(b, j)) return rc;
    *bytes = j.bytes;
    return RTM_HIP_OK;
}

int rtm_hip_batch_submit(rtm_hip_batch* b, const rtm_hip_job* job, long long* id) {
    if (int rc = check(b)) return rc;
    if (!job || !job->v || !job->shot.rx || !job->shot.rz || !job->shot.traces)
        return b->fail(RTM_HIP_ERR_ARG, "job needs a velocity model, receivers and traces");
    auto j = std::make_unique<Job>();
    j->grid = job->grid;
    j->params = job->params;
    j->nt = job->nt;
    j->dt = job->dt;
    j->sx = job->shot.sx;
    j->sz = job->shot.sz;
    const int nrec = std::max(job->shot.nrec, 0);
    j->rx.assign(job->shot.rx, job->shot.rx + nrec);
    j->rz.assign(job->shot.rz, job->shot.rz + nrec);
    j->trace_nt = job->shot.nt;
    j->traces.assign(job->shot.traces, job->shot.traces + (size_t)nrec * std::max(job->shot.nt, 0));
    j->stack_ix0 = job->stack_ix0;
    j->stack_iz0 = job->stack_iz0;
    if (int rc = derive(b, *j)) return rc;
    if (j->stack_ix0 >= 0) {
        if (!b->d_stack_i) return b->fail(RTM_HIP_ERR_ARG, "job has a stack position but no stack is set");
        if (j->stack_iz0 < 0 || j->stack_ix0 + j->grid.nx > b->stack_nx || j->stack_iz0 + j->grid.nz > b->stack_nz)
            return b->fail(RTM_HIP_ERR_ARG, "job image does not fit in the stack at its position");
    }
    j->v.assign(job->v, job->v + j->n_int);
    if (j->bytes > b->budget)
        return b->fail(RTM_HIP_ERR_MEMORY, "job needs %.2f GiB, more than the %.2f GiB budget", j->bytes / 1073741824.0,
                       b->budget / 1073741824.0);
    j->id = b->next_id++;
    j->t_submit = now_s();
    b->queue.push_back(j.get());
    if (id) *id = j->id;
    b->jobs[j->id] = std::move(j);
    b->stats.jobs_submitted += 1;
    return RTM_HIP_OK;
}

int rtm_hip_batch_run(rtm_hip_batch* b, int max_ticks, int until_done) {
    if (int rc = check(b)) return rc;
    (void)hipSetDevice(b->device);
    long long ran = 0;
    for (;;) {
        if (int rc = admit_from_queue(b)) return rc;
        if (b->active.empty()) break;
        if (b->dirty)
            if (int rc = rebuild(b)) return rc;
        if (int rc = autotune(b)) return rc;
        if (b->dirty)
            if (int rc = rebuild(b)) return rc;
        // run until the next job finishes (every job moves one step per tick)
        int n = b->active.front()->steps_left;
        for (Job* j : b->active) n = std::min(n, j->steps_left);
        if (max_ticks > 0) n = (int)std::min<long long>(n, max_ticks - ran);
        if (n <= 0) break;
        const double t0 = now_s();
        if (int rc = run_ticks(b, n)) return rc;
        ran += n;
        if (int rc = retire_finished(b)) return rc;
        b->stats.run_seconds += now_s() - t0;
        if (!until_done && max_ticks > 0 && ran >= max_ticks) break;
        if (!until_done && max_ticks <= 0) break;
    }
    hipError_t e = hipStreamSynchronize(b->stream);
    return e == hipSuccess ? RTM_HIP_OK : b->hip_fail(e, "batch run");
}

int rtm_hip_batch_job_state(rtm_hip_batch* b, long long id, int* state, double* latency_s) {
    if (int rc = check(b)) return rc;
    auto it = b->jobs.find(id);
    if (it == b->jobs.end()) return b->fail(RTM_HIP_ERR_ARG, "unknown job %lld", id);
    if (state) *state = it->second->state;
    if (latency_s)
        *latency_s = it->second->state == RTM_HIP_JOB_DONE ? it->second->t_done - it->second->t_submit : -1.0;
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
    (void)hipSetDevice(b->device);
    (void)hipStreamSynchronize(b->stream);
    if (b->d_stack_i) (void)hipFree(b->d_stack_i);
    if (b->d_stack_s) (void)hipFree(b->d_stack_s);
    b->d_stack_i = b->d_stack_s = nullptr;
    const size_t n = (size_t)nx * nz;
    hipError_t e = hipMalloc(&b->d_stack_i, n * sizeof(double));
    if (e == hipSuccess) e = hipMalloc(&b->d_stack_s, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(b->d_stack_i, 0, n * sizeof(double));
    if (e == hipSuccess) e = hipMemset(b->d_stack_s, 0, n * sizeof(double));
    if (e != hipSuccess) return b->hip_fail(e, "FP64 stack");
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
    out->jobs_queued = (int)b->queue.size();
    out->bytes_in_use = b->used;
    out->bytes_budget = b->budget;
    int k = 0;
    for (auto& [half, g] : b->groups) {
        if (k >= 6) break;
        out->group_half[k] = half;
        out->group_tile_x[k] = g.tile_x;
        out->group_dpp[k] = g.dpp ? 1 : 0;
        ++k;
    }
    out->ngroups = k;
    return RTM_HIP_OK;
}

}  // extern "C"
