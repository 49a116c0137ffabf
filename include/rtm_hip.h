/*
 * rtm_hip.h -- C ABI of the MI355X (gfx950) 2D acoustic RTM engine.
 *
 * The reference has no foreign-function layer: its engines are C++ classes
 * behind rtm::RTMEngine (include/rtm_engine.hpp:20-83) selected by
 * rtm::make_engine (include/rtm_factory.hpp:19-23).  This header flattens that
 * interface into plain C so that any host (the reference's own C++ factory, the
 * C++ HipRTM wrapper in include/rtm/rtm_hip_engine.hpp, ctypes) can drive the
 * HIP kernels.  Each entry point names the reference method it replaces.
 *
 * Conventions
 *   - Every array is depth-fastest: interior index ix*nz+iz, as in the
 *     reference (include/rtm_types.hpp:10-28).  Traces are trace-major,
 *     traces[ir*nt_stride + it] (include/rtm_types.hpp:54).
 *   - Host buffers are caller-owned and only read or written during the call.
 *   - Functions return RTM_HIP_OK (0) or a negative status; the message is
 *     available from rtm_hip_last_error(handle).  The C++ wrapper turns a
 *     negative status into std::runtime_error, which is the reference's error
 *     behaviour (src/cuda/cuda_check.hpp:7-14).
 *   - A handle owns one HIP stream on one device, plus every table, wavefield,
 *     snapshot and image buffer of its engine.  Handles are not thread-safe;
 *     distinct handles may be used from distinct threads.
 */
#ifndef RTM_HIP_H
#define RTM_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTM_HIP_OK            0
#define RTM_HIP_ERR_ARG      -1   /* invalid argument or state           */
#define RTM_HIP_ERR_DEVICE   -2   /* HIP runtime error (incl. no device) */
#define RTM_HIP_ERR_MEMORY   -3   /* device or host allocation failed    */

/* Step-kernel families (rtm_hip_set_step_kernel). */
#define RTM_HIP_STEP_LDS_TILE  0
#define RTM_HIP_STEP_MARCH     1
#define RTM_HIP_STEP_MARCH_DPP 2
#define RTM_HIP_STEP_STREAM    3

/* Snapshot storage formats (rtm_hip_setup's params->snapshot_format). */
#define RTM_HIP_SNAP_F32   0      /* the reference's format (src/rtm/forward.cpp:83-88) */
#define RTM_HIP_SNAP_F16   1      /* IEEE half + one FP32 scale per snapshot            */
#define RTM_HIP_SNAP_BF16  2      /* bfloat16 + one FP32 scale per snapshot             */

/* Cross-shot image stack (rtm_hip_setup's params->image_stack). */
#define RTM_HIP_STACK_F32  0      /* FP32 running sum over shots, the reference's (src/rtm/rtm.cpp:92-101) */
#define RTM_HIP_STACK_F64  1      /* per-shot FP32 image added into an FP64 stack after each shot      */

/* rtm::Grid, interior part (include/rtm_types.hpp:16-29). */
typedef struct {
    int   nx, nz;
    float dx, dz;
    float ox, oz;
} rtm_hip_grid;

/* rtm::RTMParams (include/rtm_types.hpp:33-40) plus storage precisions. */
typedef struct {
    int   nb;
    int   order;
    float sponge_alpha;
    int   store_interval;
    float f0;
    int   verbose;
    int   snapshot_format;       /* RTM_HIP_SNAP_*; F32 reproduces the reference       */
    int   image_stack;           /* RTM_HIP_STACK_*; F32 reproduces the reference      */
} rtm_hip_params;

/* rtm::ShotRecord (include/rtm_types.hpp:49-57). */
typedef struct {
    float        sx, sz;         /* source position, metres             */
    int          nrec;
    const float* rx;             /* receiver positions, metres (nrec)   */
    const float* rz;
    int          nt;             /* samples per trace = trace stride    */
    float        dt;
    const float* traces;         /* nrec*nt trace-major, or NULL        */
} rtm_hip_shot;

/* rtm::StageTimes (include/rtm_engine.hpp:8-13), seconds, accumulated since
 * setup.  imaging stays 0: imaging is fused into the backward step kernel. */
typedef struct {
    double io, forward, backward, imaging, total, h2d, kernel, d2h;
} rtm_hip_times;

typedef struct {
    int    nx, nz, nb, half, nxe, nze;
    int    ldz;                  /* device row pitch of extended fields  */
    int    nt, store_interval, num_snapshots;
    size_t snapshot_bytes;       /* device bytes of the snapshot store   */
    size_t device_bytes;         /* all device buffers of this handle    */
    int    snapshot_format;
    int    image_stack;
    int    step_family[2];       /* RTM_HIP_STEP_* of the forward/backward step kernel */
    int    tile_x[2];            /* their tile widths along x            */
    int    autotuned[2];         /* 1 when chosen by the autotuner       */
    int    waves_per_cu[2];      /* resident wavefronts per CU the launch is capped at (0: no cap) */
    char   device_name[128];
} rtm_hip_info;

/* Sampled step-kernel timing (HIP events around every `every`-th launch on
 * the handle's stream).  kind 0 = forward step kernel, 1 = backward. */
typedef struct {
    long long launches[2];       /* all step launches since reset        */
    long long sampled[2];        /* launches bracketed by events         */
    double    sampled_ms[2];     /* summed event time of sampled launches */
} rtm_hip_kernel_stats;

typedef struct rtm_hip_engine rtm_hip_engine;

/* Number of visible HIP devices (0 when none); never fails. */
int         rtm_hip_device_count(void);

/* Replaces CUDARTM's constructor + cudaSetDevice(0) (src/cuda/rtm_cuda.cu:30-33):
 * binds a handle to `device` and creates its stream.  Fails cleanly with
 * RTM_HIP_ERR_DEVICE when there is no such device. */
int         rtm_hip_create(int device, rtm_hip_engine** out);

/* Replaces CUDARTM::~CUDARTM (src/cuda/rtm_cuda.cu:124-138). NULL is a no-op. */
void        rtm_hip_destroy(rtm_hip_engine* h);

/* Message of the last failure on h (h == NULL: of the last failed create or
 * of a call made with a NULL handle).  Never NULL. */
const char* rtm_hip_last_error(const rtm_hip_engine* h);

/* Replaces RTMEngine::setup (include/rtm_engine.hpp:25-27; CPUReferenceRTM::setup
 * src/rtm/rtm_cpu.cpp:9-88; CUDARTM::setup src/cuda/rtm_cuda.cu:30-116).
 * v: interior velocity nx*nz, m/s.  Builds (v dt)^2, the Cerjan tapers, the
 * delayed Ricker wavelet and the stencil coefficients on the host with the
 * reference's float formulas, uploads them, and allocates the wavefields,
 * the snapshot store (num_snapshots = (nt-1)/store_interval + 1) and the
 * image/illumination accumulators (zeroed). */
int         rtm_hip_setup(rtm_hip_engine* h, const rtm_hip_grid* grid, const float* v,
                          const rtm_hip_params* params, int nt, float dt);

int         rtm_hip_get_info(const rtm_hip_engine* h, rtm_hip_info* out);

/* Replaces RTMEngine::forward_propagation (include/rtm_engine.hpp:34-36;
 * src/rtm/forward.cpp:97-125; src/cuda/propagation_cuda.cu:160-230).
 * Propagates the source wavefield and fills the device snapshot store.
 * snapshots_out: NULL (device-resident, as the CUDA engine) or a host buffer
 * of num_snapshots*nx*nz floats that receives the stored snapshots
 * (dequantised when the format is F16/BF16).  recorded: NULL or nrec*nt
 * floats receiving the modelled traces, trace-major with stride nt. */
int         rtm_hip_forward(rtm_hip_engine* h, const rtm_hip_shot* shot,
                            float* snapshots_out, float* recorded);

/* Replaces RTMEngine::backward_propagation (include/rtm_engine.hpp:39-42;
 * src/rtm/backward.cpp:40-75): back-propagates shot->traces (stride
 * shot->nt, which may exceed the engine's nt, as with the reference's --nt)
 * and accumulates I += p_s*p_r, S += p_s^2 into the device image on every
 * stored step, using the snapshots of the preceding rtm_hip_forward. */
int         rtm_hip_backward(rtm_hip_engine* h, const rtm_hip_shot* shot);

/* Zero / download the device image and illumination (nx*nz floats each).
 * Replaces the device image handling of CUDARTM::migrate
 * (src/cuda/rtm_cuda.cu:222-244). */
int         rtm_hip_image_zero(rtm_hip_engine* h);
int         rtm_hip_image_get(rtm_hip_engine* h, float* image, float* illum);
/* The FP64 stacks of an engine set up with RTM_HIP_STACK_F64 (no reference
 * counterpart: the reference stacks in FP32). */
int         rtm_hip_image_get_f64(rtm_hip_engine* h, double* image, double* illum);

/* Post-processing of the resident image on the device.  Replaces
 * illumination_compensation and laplacian_filter (src/rtm/imaging.cpp:48-75),
 * which the reference's rtm applies on the host to copies of the stacked
 * image for its _illum.bin and _filtered.bin outputs (src/main.cpp:211-222):
 *   RTM_HIP_POST_ILLUM      out = I / (S + eps_relative*max(S) + 1e-30f)
 *   RTM_HIP_POST_LAPLACIAN  out = Laplacian high-pass of I (of the compensated
 *                           image when both flags are set), 0 on the outer ring
 * out receives nx*nz floats; the device image and illumination are left
 * unchanged.  With an FP64 stack the source is the stack rounded to float, as
 * rtm_hip_image_get returns it.  Bit-identical to the reference's host code. */
#define RTM_HIP_POST_ILLUM      1
#define RTM_HIP_POST_LAPLACIAN  2
int         rtm_hip_image_postprocess(rtm_hip_engine* h, int flags, float eps_relative, float* out);

/* rtm_hip_image_postprocess with no allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and bits; d_out is a device buffer of nx*nz floats the caller
 * owns; the only scratch (512 partial maxima, 2 KiB) is a static device array
 * of the library, one per GPU, shared by every call on that GPU: calls that
 * overlap on one GPU race.  Calls no hipMalloc, hipFree or allocating routine;
 * returns when the result is in d_out. */
int         rtm_hip_image_postprocess_noalloc(rtm_hip_engine* h, int flags, float eps_relative, float* d_out);

/* Replaces RTMEngine::migrate (include/rtm_engine.hpp:45-47; src/rtm/rtm.cpp:83-103;
 * src/cuda/rtm_cuda.cu:214-250): zero the image, forward + backward every shot
 * in order (FP32 running sum on the device), download image and illum. */
int         rtm_hip_migrate(rtm_hip_engine* h, int nshots, const rtm_hip_shot* shots,
                            float* image, float* illum);

/* Replaces RTMEngine::model_shot (include/rtm_engine.hpp:50; src/rtm/rtm.cpp:105-111):
 * forward modelling with trace recording; traces_out nrec*nt (engine nt). */
int         rtm_hip_model_shot(rtm_hip_engine* h, const rtm_hip_shot* shot, float* traces_out);

/* Replaces RTMEngine::times (include/rtm_engine.hpp:75). */
int         rtm_hip_get_times(const rtm_hip_engine* h, rtm_hip_times* out);
int         rtm_hip_reset_times(rtm_hip_engine* h);

/* Step-kernel timing for the roofline figure.  every > 0: HIP events around
 * every `every`-th step launch (each such launch also carries the events'
 * end-of-kernel overhead); every < 0: one event pair around all step
 * launches of each forward/backward loop, counted as nt launches (the mean
 * then includes the short gaps between back-to-back launches); 0: off.
 * No reference counterpart (the reference never filled times.kernel,
 * include/rtm_engine.hpp:11). */
int         rtm_hip_set_kernel_timing(rtm_hip_engine* h, int every);
int         rtm_hip_get_kernel_stats(const rtm_hip_engine* h, rtm_hip_kernel_stats* out);

/* Step-kernel mapping: family RTM_HIP_STEP_MARCH / RTM_HIP_STEP_MARCH_DPP (one
 * wave64 per depth strip, all loads up front; depth neighbours by ds_bpermute
 * or by DPP lane shifts), RTM_HIP_STEP_STREAM (the strip marches along its
 * columns with p^n in a sliding register window) or RTM_HIP_STEP_LDS_TILE
 * (LDS-staged 64-cell tiles), or -1 for the built-in choice; tile_x = output
 * columns per tile (4, 8 or 16; 8, 16 or 32 for the stream kernel), 0 for the
 * built-in choice.  The built-in choice is an autotuner that times
 * every family and width on the engine's grid before the first propagation
 * of each direction after setup.  Results are identical for every choice;
 * takes effect at the next forward/backward.  No reference counterpart (the CUDA
 * block shape is fixed, src/cuda/propagation_cuda.cu:96-97). */
int         rtm_hip_set_step_kernel(rtm_hip_engine* h, int family, int tile_x);

/* Replaces the reference's GPU bandwidth probe (src/tools/bandwidth_probe_cuda.cu:10-66):
 * copies `bytes` between two device buffers `repeats` times (after one
 * warm-up) and returns the best (read + write) rate in GB/s.  Allocates and
 * frees its own buffers on `device`. */
int         rtm_hip_bandwidth_probe(int device, size_t bytes, int repeats, double* best_gbs);

/* Replaces rtm::mute_direct_wave (include/rtm_io.hpp:9; src/io/seismic_io.cpp:18-35),
 * host-only: traces (nrec*nt, stride shot->nt) muted in place. */
void        rtm_hip_mute_direct_wave(const rtm_hip_shot* shot, float* traces,
                                     float v_surface, float t_pad, float t_taper);

/* ------------------------------------------------------------------------
 * Batched and served migration (rtm_hip_batch_*).
 *
 * The reference's engine binds one model, one RTMParams and one time axis per
 * instance (include/rtm_engine.hpp:25-27) and migrates shots one after the
 * other (src/rtm/rtm.cpp:83-103).  A batch handle instead takes independent
 * shot-migration jobs, each with its own grid, velocity, order, store
 * interval and nt, and advances every job on the device by one time step per
 * tick, all jobs of one stencil order in one launch (a tile table spans the
 * jobs), forward and backward phases mixed.  Jobs are admitted from a FIFO
 * queue when their device footprint fits the memory budget; finished jobs
 * free their memory and queued jobs take it (continuous batching).  A job's
 * image is bit-identical to its image from rtm_hip_migrate of that one shot
 * on a plain handle with the same snapshot format.  No reference counterpart.
 * ------------------------------------------------------------------------ */
#define RTM_HIP_JOB_QUEUED   0
#define RTM_HIP_JOB_RUNNING  1
#define RTM_HIP_JOB_DONE     2
#define RTM_HIP_JOB_FAILED   3

typedef struct {
    size_t memory_budget;        /* device bytes for jobs; 0 = 90% of free memory    */
    int    max_active;           /* cap on jobs on the device at once; 0 = no cap    */
    int    use_graph;            /* 1: replay captured runs of ticks (hipGraph) while ticks
                                    are small enough that host launches would pace them */
    int    kernel_family;        /* 0: autotuned per stencil order and load; else RTM_HIP_STEP_MARCH,
                                    RTM_HIP_STEP_MARCH_DPP or RTM_HIP_STEP_STREAM      */
    int    tile_x;               /* with kernel_family: its tile width (0: default)  */
    int    merge_groups;         /* 1: one launch per tick for jobs of every stencil order and
                                    snapshot width (DPP march, 2-4 column tiles); 0: one launch
                                    per group, the groups on concurrent streams */
} rtm_hip_batch_config;

/* One shot migration: the arguments of setup (grid, velocity, params, nt, dt)
 * and of migrate for one shot.  params->snapshot_format applies per job;
 * params->image_stack is ignored (see stack_ix0).  shot->traces is required,
 * with shot->nt >= nt samples per trace.  stack_ix0/stack_iz0 >= 0 add the
 * finished job's image and illumination into the handle's FP64 stack at
 * that offset (rtm_hip_batch_set_stack); -1 leaves the stack alone. */
typedef struct {
    rtm_hip_grid   grid;
    const float*   v;            /* nx*nz interior velocity */
    rtm_hip_params params;
    int            nt;
    float          dt;
    rtm_hip_shot   shot;
    int            stack_ix0, stack_iz0;
} rtm_hip_job;

typedef struct {
    long long jobs_submitted, jobs_done;
    int       jobs_active, jobs_queued;
    long long ticks;             /* ticks run (each advances every active job one step; a
                                    job takes 2*nt ticks, plus one with 16-bit snapshots) */
    long long rebuilds;          /* launch-table rebuilds (job set changes)         */
    double    cell_steps;        /* 2*nt*nxe*nze summed over finished jobs          */
    double    run_seconds;       /* wall time inside rtm_hip_batch_run, of which: */
    double    admit_seconds;     /*   admission (allocation, tables, upload), also
                                    ahead of a change while the device runs ticks */
    double    tune_seconds;      /*   kernel autotuning                            */
    double    retire_seconds;    /*   image download, stack update, freeing        */
    double    busy_latency_s;    /* sum of submit-to-done times of finished jobs   */
    int       models_resident;   /* registered models (rtm_hip_batch_add_model)     */
    size_t    bytes_in_use, bytes_budget;
    int       ngroups;           /* launch groups: (half-width, snapshot width)    */
    int       group_half[12], group_tile_x[12], group_family[12]; /* RTM_HIP_STEP_* */
    int       group_waves_per_cu[12];                             /* 0: no cap */
    int       group_snap16[12];  /* 1: the group's jobs store 16-bit snapshots     */
    int       tunes;             /* autotune searches (one per group and load class) */
    double    rebuild_seconds;   /* launch-table rebuilds (inside run_seconds)      */
    double    capture_seconds;   /* hipGraph capture and instantiation (inside run) */
    long long graphs_captured;   /* captured runs of ticks, one per launch pattern  */
    long long step_launches;     /* step launches (one per tick and group; one per tick
                                    and job on per-job streams)                    */
    long long lean_launches;     /* of which ran the lean build (no snapshot work)  */
    int       merged_launch;     /* 1: the groups present run in one launch per tick */
    long long concurrent_ticks;  /* ticks run with each job on the per-shot kernels on a
                                    stream of its own (2-3 jobs, one filling the chip) */
    long long staged_admissions; /* jobs admitted ahead, while the ticks before the job
                                    set's next change ran (counted queued until then) */
} rtm_hip_batch_stats;

typedef struct rtm_hip_batch rtm_hip_batch;

/* Creates a batch handle on `device` (cfg NULL: defaults). */
int         rtm_hip_batch_create(int device, const rtm_hip_batch_config* cfg, rtm_hip_batch** out);
void        rtm_hip_batch_destroy(rtm_hip_batch* b);
const char* rtm_hip_batch_last_error(const rtm_hip_batch* b);

/* Replaces RTMEngine::setup (include/rtm_engine.hpp:25-27) for the batch:
 * validates the model, parameters and time axis as setup does, builds the
 * tables on the host with the reference's float formulas and uploads them
 * now.  The model stays resident (and counts against the budget) until
 * removed, and serves every shot submitted on it; params->snapshot_format
 * applies to those shots. */
int         rtm_hip_batch_add_model(rtm_hip_batch* b, const rtm_hip_grid* grid, const float* v,
                                    const rtm_hip_params* params, int nt, float dt, int* model_id);
/* Forgets the model; its tables are freed once no queued or active job uses it. */
int         rtm_hip_batch_remove_model(rtm_hip_batch* b, int model_id);

/* Queues one shot on a resident model (the per-shot body of RTMEngine::migrate,
 * src/rtm/rtm.cpp:88-101); stack position as in rtm_hip_job. */
int         rtm_hip_batch_submit_shot(rtm_hip_batch* b, int model_id, const rtm_hip_shot* shot,
                                      int stack_ix0, int stack_iz0, long long* id);

/* Reserves `bytes` of device memory (0: the whole remaining budget) in one
 * allocation, from which every later job and model buffer is placed; the
 * budget becomes what is in use plus the reservation.  Admission then costs
 * no device allocation (a large allocation also clears its memory, which is
 * the dominant cost of admitting a job otherwise).  Only before the first
 * submission. */
int         rtm_hip_batch_reserve(rtm_hip_batch* b, size_t bytes);

/* Device bytes admission needs for the job (its own buffers plus the tables
 * of its private model). */
int         rtm_hip_batch_job_bytes(rtm_hip_batch* b, const rtm_hip_job* job, size_t* bytes);

/* Validates and queues a job with a private model (inputs copied; the
 * caller's buffers are free on return) and returns its id.  Nothing runs
 * until rtm_hip_batch_run. */
int         rtm_hip_batch_submit(rtm_hip_batch* b, const rtm_hip_job* job, long long* id);

/* Admits queued jobs and runs ticks.  until_done = 1: until no job is queued
 * or active (or max_ticks > 0 ticks have run).  until_done = 0: at most
 * max_ticks ticks (max_ticks <= 0: until the next job finishes).  Jobs that
 * finish are retired (image downloaded, stack updated, memory freed).  A call
 * with until_done = 0 that leaves jobs running returns with its last ticks
 * still queued on the device, once the ticks the previous such call left
 * are done, so a serving loop keeps the device busy between calls. */
int         rtm_hip_batch_run(rtm_hip_batch* b, int max_ticks, int until_done);

/* State (RTM_HIP_JOB_*) and submit-to-done latency (-1 until done). */
int         rtm_hip_batch_job_state(rtm_hip_batch* b, long long id, int* state, double* latency_s);

/* Image and illumination (nx*nz floats each, either may be NULL) of a
 * finished job; the job is forgotten afterwards. */
int         rtm_hip_batch_result(rtm_hip_batch* b, long long id, float* image, float* illum);

/* Allocates (and zeroes) an nx*nz FP64 image/illumination stack; refused
 * while queued or active jobs have a stack position. */
int         rtm_hip_batch_set_stack(rtm_hip_batch* b, int nx, int nz);
int         rtm_hip_batch_get_stack(rtm_hip_batch* b, double* image, double* illum);

int         rtm_hip_batch_get_stats(rtm_hip_batch* b, rtm_hip_batch_stats* out);

#ifdef __cplusplus
}
#endif
#endif /* RTM_HIP_H */
