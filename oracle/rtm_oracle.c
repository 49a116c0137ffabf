/*
 * rtm_oracle.c -- CPU restatement of the reference RTM path (FP32 + FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see rtm_oracle.h).  Built at -O2 without FP
 * contraction, like the reference's own build (CMakeLists.txt:15-17), so the
 * FP32 instance reproduces the reference's bits.  The FP32 "cpu" instance is
 * single-threaded like the reference's; the FP64 instance runs its cell loops
 * on every host core (OpenMP), which changes no result.
 */
#define _GNU_SOURCE
#include "rtm_oracle_internal.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static __thread char g_err[512];

void orc_set_error(const char* msg) {
    snprintf(g_err, sizeof g_err, "%s", msg);
}

const char* orc_last_error(void) { return g_err; }

double orc_now(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ---- numerics, src/rtm/rtm.cpp:14-70 ------------------------------------ */

int orc_stencil_half(int order) {
    switch (order) {
        case 2: case 4: case 6: case 8: case 10: case 12: return order / 2;
        default: return -1;
    }
}

/* Taylor second-derivative coefficients as exact fractions
 * (src/rtm/rtm.cpp:14-23).  The FP32 set is written with float literals and
 * float divisions, as the reference's tables are; the FP64 set divides in
 * double. */
int orc_stencil_coefficients_f32(int order, float* out) {
    switch (order) {
        case 2:  out[0] = -2.0f; out[1] = 1.0f; break;
        case 4:  out[0] = -2.5f; out[1] = 4.0f / 3.0f; out[2] = -1.0f / 12.0f; break;
        case 6:  out[0] = -49.0f / 18.0f; out[1] = 1.5f; out[2] = -0.15f;
                 out[3] = 1.0f / 90.0f; break;
        case 8:  out[0] = -205.0f / 72.0f; out[1] = 8.0f / 5.0f; out[2] = -1.0f / 5.0f;
                 out[3] = 8.0f / 315.0f; out[4] = -1.0f / 560.0f; break;
        case 10: out[0] = -5269.0f / 1800.0f; out[1] = 5.0f / 3.0f; out[2] = -5.0f / 21.0f;
                 out[3] = 5.0f / 126.0f; out[4] = -5.0f / 1008.0f; out[5] = 1.0f / 3150.0f;
                 break;
        case 12: out[0] = -5369.0f / 1800.0f; out[1] = 12.0f / 7.0f; out[2] = -15.0f / 56.0f;
                 out[3] = 10.0f / 189.0f; out[4] = -1.0f / 112.0f; out[5] = 2.0f / 1925.0f;
                 out[6] = -1.0f / 16632.0f; break;
        default: return -1;
    }
    return order / 2;
}

int orc_stencil_coefficients_f64(int order, double* out) {
    switch (order) {
        case 2:  out[0] = -2.0; out[1] = 1.0; break;
        case 4:  out[0] = -2.5; out[1] = 4.0 / 3.0; out[2] = -1.0 / 12.0; break;
        case 6:  out[0] = -49.0 / 18.0; out[1] = 1.5; out[2] = -0.15;
                 out[3] = 1.0 / 90.0; break;
        case 8:  out[0] = -205.0 / 72.0; out[1] = 8.0 / 5.0; out[2] = -1.0 / 5.0;
                 out[3] = 8.0 / 315.0; out[4] = -1.0 / 560.0; break;
        case 10: out[0] = -5269.0 / 1800.0; out[1] = 5.0 / 3.0; out[2] = -5.0 / 21.0;
                 out[3] = 5.0 / 126.0; out[4] = -5.0 / 1008.0; out[5] = 1.0 / 3150.0;
                 break;
        case 12: out[0] = -5369.0 / 1800.0; out[1] = 12.0 / 7.0; out[2] = -15.0 / 56.0;
                 out[3] = 10.0 / 189.0; out[4] = -1.0 / 112.0; out[5] = 2.0 / 1925.0;
                 out[6] = -1.0 / 16632.0; break;
        default: return -1;
    }
    return order / 2;
}

/* |c0| + 2 sum |ck| in float (src/rtm/rtm.cpp:43-49). */
float orc_stencil_abs_sum(int order) {
    float c[ORC_MAX_HALF + 1];
    const int h = orc_stencil_coefficients_f32(order, c);
    if (h < 0) return -1.0f;
    float s = fabsf(c[0]);
    for (int k = 1; k <= h; ++k) s += 2.0f * fabsf(c[k]);
    return s;
}

/* src/rtm/rtm.cpp:52-55 */
float orc_ricker(float t, float f0) {
    const float a = (float)(M_PI * M_PI) * f0 * f0 * t * t;
    return (1.0f - 2.0f * a) * expf(-a);
}

/* src/rtm/rtm.cpp:59-62 */
float orc_max_stable_dt(float vmax, float dx, float dz, int order) {
    const float S = orc_stencil_abs_sum(order);
    return 2.0f / (vmax * sqrtf(S * (1.0f / (dx * dx) + 1.0f / (dz * dz))));
}

/* src/rtm/rtm.cpp:66-70 */
float orc_points_per_wavelength(float vmin, float f0, float dx, float dz) {
    const float fmax = 2.5f * f0;
    const float lambda_min = vmin / fmax;
    return lambda_min / (dx > dz ? dx : dz);
}

/* Direct-arrival mute, src/io/seismic_io.cpp:18-35. */
void orc_mute_direct_wave(orc_shot* s, float* traces, float v_surface, float t_pad,
                          float t_taper) {
    if (v_surface <= 0.0f) return;
    for (int ir = 0; ir < s->nrec; ++ir) {
        const float ddx = s->rx[ir] - s->sx;
        const float ddz = s->rz[ir] - s->sz;
        const float off = sqrtf(ddx * ddx + ddz * ddz);
        const float t0 = off / v_surface + t_pad;
        const int i0 = (int)(t0 / s->dt);
        int ntap = (int)(t_taper / s->dt);
        if (ntap < 1) ntap = 1;
        const int i1 = i0 + ntap;
        float* tr = traces + (size_t)ir * s->nt;
        for (int it = 0; it < s->nt && it < i1; ++it) {
            if (it <= i0) { tr[it] = 0.0f; continue; }
            const float u = (float)(it - i0) / (float)(i1 - i0);
            tr[it] *= 0.5f * (1.0f - cosf((float)M_PI * u));
        }
    }
}

/* ---- engine lifecycle ---------------------------------------------------- */

orc_engine* orc_create(int precision, int variant) {
    if (precision != 32 && precision != 64) { orc_set_error("precision must be 32 or 64"); return NULL; }
    if (variant != 0 && variant != 1) { orc_set_error("variant must be 0 (cpu) or 1 (cpu-opt)"); return NULL; }
    if (variant == 1 && precision != 32) { orc_set_error("cpu-opt exists in FP32 only"); return NULL; }
    orc_engine* e = (orc_engine*)calloc(1, sizeof(orc_engine));
    if (!e) { orc_set_error("out of host memory"); return NULL; }
    e->precision = precision;
    e->variant = variant;
    return e;
}

static void free_tables(orc_engine* e) {
    free(e->vdt2); free(e->sponge); free(e->wavelet);
    for (int b = 0; b < 3; ++b) { free(e->buf[b]); e->buf[b] = NULL; }
    e->vdt2 = e->sponge = e->wavelet = NULL;
}

void orc_destroy(orc_engine* e) {
    if (!e) return;
    free_tables(e);
    free(e->rec);
    free(e);
}

int orc_num_snapshots(const orc_engine* e) { return e ? e->nsnap : 0; }

int orc_num_threads(const orc_engine* e) {
    return (e && e->variant == 1) ? orc_opt_num_threads() : 1;
}

void orc_times(const orc_engine* e, double* fwd, double* bwd, double* img) {
    *fwd = e->t_fwd; *bwd = e->t_bwd; *img = e->t_img;
}

/* Nearest-grid-point mapping with clamping, src/rtm/rtm_cpu.cpp:99-123.
 * Done in float for both precisions so that FP32 and FP64 runs share the
 * same cells. */
static size_t to_index(const orc_engine* e, float x, float z, int* nclamped) {
    const int ix0 = (int)lroundf((x - e->ox) / e->dx);
    const int iz0 = (int)lroundf((z - e->oz) / e->dz);
    const int ix = ix0 < 0 ? 0 : (ix0 > e->nx - 1 ? e->nx - 1 : ix0);
    const int iz = iz0 < 0 ? 0 : (iz0 > e->nz - 1 ? e->nz - 1 : iz0);
    if (ix != ix0 || iz != iz0) ++*nclamped;
    return (size_t)(ix + e->nb) * e->nze + (size_t)(iz + e->nb);
}

int orc_map_geometry_internal(orc_engine* e, const orc_shot* s) {
    int nclamped = 0;
    e->src = to_index(e, s->sx, s->sz, &nclamped);
    if (s->nrec > e->nrec_cap) {
        free(e->rec);
        e->rec = (size_t*)malloc((size_t)s->nrec * sizeof(size_t));
        if (!e->rec) { e->nrec_cap = 0; orc_set_error("out of host memory"); return -1; }
        e->nrec_cap = s->nrec;
    }
    e->nrec = s->nrec;
    for (int ir = 0; ir < s->nrec; ++ir)
        e->rec[ir] = to_index(e, s->rx[ir], s->rz[ir], &nclamped);
    if (nclamped > 0 && e->verbose)
        fprintf(stderr, "oracle: %d of %d source/receiver positions clamped to the model edge\n",
                nclamped, s->nrec + 1);
    return 0;
}

int orc_map_geometry(orc_engine* e, const orc_shot* s, long long* src_out, long long* rec_out) {
    if (!e || e->nze == 0) { orc_set_error("engine not set up"); return -1; }
    if (orc_map_geometry_internal(e, s) != 0) return -1;
    *src_out = (long long)e->src;
    for (int ir = 0; ir < e->nrec; ++ir) rec_out[ir] = (long long)e->rec[ir];
    return 0;
}

/* ---- precision-generic body ---------------------------------------------- */

#define REAL float
#define SFX _f32
#define PREC 32
#define ONE 1.0f
#define TWO 2.0f
#define PI_SQ ((float)(M_PI * M_PI))
#define T0_NUM 1.2f
#define EXPR expf
#define CF(name) (e->name##_f)
#define ORC_OPT_STEP(e, a, b, c) orc_opt_fused_step((e), (a), (b), (c))
#define SAVE_SNAPSHOT(e, p, d) \
    do { if ((e)->variant == 1) orc_opt_save_snapshot((e), (p), (d)); else save_snapshot__f32((e), (p), (d)); } while (0)
#define IMAGING(e, f, b, I, S) \
    do { if ((e)->variant == 1) orc_opt_imaging((e), (f), (b), (I), (S)); else imaging__f32((e), (f), (b), (I), (S)); } while (0)
#include "rtm_oracle_core.inc"
#undef REAL
#undef SFX
#undef PREC
#undef ONE
#undef TWO
#undef PI_SQ
#undef T0_NUM
#undef EXPR
#undef CF
#undef ORC_OPT_STEP
#undef SAVE_SNAPSHOT
#undef IMAGING

#define REAL double
#define SFX _f64
#define PREC 64
#define ONE 1.0
#define TWO 2.0
#define PI_SQ (M_PI * M_PI)
#define T0_NUM ((double)1.2f)
#define EXPR exp
#define CF(name) (e->name##_d)
/* cpu-opt exists in FP32 only (orc_create refuses variant 1 here). */
#define ORC_OPT_STEP(e, a, b, c) fd_step__f64((e), (a), (b), (c))
#define SAVE_SNAPSHOT(e, p, d) save_snapshot__f64((e), (p), (d))
#define IMAGING(e, f, b, I, S) imaging__f64((e), (f), (b), (I), (S))
#include "rtm_oracle_core.inc"

/* ---- setup ---------------------------------------------------------------- */

/* CPUReferenceRTM::setup validation and sizes, src/rtm/rtm_cpu.cpp:9-24. */
int orc_setup(orc_engine* e, const orc_grid* g, const float* v, const orc_params* p,
              int nt, float dt) {
    if (!e) { orc_set_error("null engine"); return -1; }
    const int half = orc_stencil_half(p->order);
    if (half < 0) { orc_set_error("unsupported spatial order (use 2, 4, 6, 8, 10 or 12)"); return -1; }
    if (p->nb < half + 1) { orc_set_error("--nb must be at least order/2 + 1"); return -1; }
    if (g->nx <= 0 || g->nz <= 0 || !v) { orc_set_error("velocity model size does not match nx*nz"); return -1; }
    if (p->store_interval < 1) { orc_set_error("--store-interval must be >= 1"); return -1; }
    if (nt < 1) { orc_set_error("nt must be >= 1"); return -1; }
    free_tables(e);
    e->nx = g->nx; e->nz = g->nz; e->dx = g->dx; e->dz = g->dz; e->ox = g->ox; e->oz = g->oz;
    e->nb = p->nb; e->half = half; e->order = p->order;
    e->nxe = g->nx + 2 * p->nb; e->nze = g->nz + 2 * p->nb;
    e->n_ext = (size_t)e->nxe * e->nze;
    e->n_int = (size_t)g->nx * g->nz;
    e->alpha = p->sponge_alpha; e->f0 = p->f0; e->store = p->store_interval;
    e->verbose = p->verbose;
    e->nt = nt; e->dt = dt;
    e->nsnap = (nt - 1) / p->store_interval + 1;
    e->t_fwd = e->t_bwd = e->t_img = 0.0;
    return e->precision == 32 ? build_tables__f32(e, v) : build_tables__f64(e, v);
}

static size_t table_len(const orc_engine* e, int which) {
    switch (which) {
        case 0: case 1: return e->n_ext;
        case 2: return (size_t)e->nt;
        case 3: return (size_t)(2 * e->half + 1);
        default: return 0;
    }
}

#define ORC_GET_TABLE(OUT_T)                                                   \
    if (!e || !e->vdt2) { orc_set_error("engine not set up"); return -1; }   \
    const size_t n = table_len(e, which);                                       \
    if (n == 0) { orc_set_error("unknown table"); return -1; }                 \
    if (which == 3) {                                                           \
        if (e->precision == 32) {                                               \
            out[0] = (OUT_T)e->c0_f;                                            \
            for (int k = 1; k <= e->half; ++k) {                                \
                out[k] = (OUT_T)e->cx_f[k]; out[e->half + k] = (OUT_T)e->cz_f[k]; } \
        } else {                                                                \
            out[0] = (OUT_T)e->c0_d;                                            \
            for (int k = 1; k <= e->half; ++k) {                                \
                out[k] = (OUT_T)e->cx_d[k]; out[e->half + k] = (OUT_T)e->cz_d[k]; } \
        }                                                                       \
        return 0;                                                               \
    }                                                                           \
    const void* src = which == 0 ? e->vdt2 : (which == 1 ? e->sponge : e->wavelet); \
    for (size_t i = 0; i < n; ++i)                                              \
        out[i] = e->precision == 32 ? (OUT_T)((const float*)src)[i]             \
                                    : (OUT_T)((const double*)src)[i];           \
    return 0;

int orc_get_table_f32(const orc_engine* e, int which, float* out) { ORC_GET_TABLE(float) }
int orc_get_table_f64(const orc_engine* e, int which, double* out) { ORC_GET_TABLE(double) }

/* RTMEngine::model_shot, src/rtm/rtm.cpp:105-111: recorded with stride nt. */
int orc_model_shot_f32(orc_engine* e, const orc_shot* s, float* traces_out) {
    memset(traces_out, 0, (size_t)s->nrec * (size_t)e->nt * sizeof(float));
    return orc_forward_f32(e, s, NULL, traces_out);
}

/* ---- post-processing, src/rtm/imaging.cpp:48-75 --------------------------- */

void orc_illumination_compensation(float* image, const float* illum, size_t n,
                                   float eps_relative) {
    float smax = 0.0f;
    for (size_t i = 0; i < n; ++i) smax = illum[i] > smax ? illum[i] : smax;
    const float eps = eps_relative * smax + 1e-30f;
    for (size_t i = 0; i < n; ++i) image[i] /= (illum[i] + eps);
}

void orc_laplacian_filter(float* image, int nx, int nz, float dx, float dz) {
    const size_t n = (size_t)nx * nz;
    float* out = (float*)calloc(n, sizeof(float));
    if (!out) return;
    const float idx2 = 1.0f / (dx * dx);
    const float idz2 = 1.0f / (dz * dz);
    for (int ix = 1; ix < nx - 1; ++ix)
        for (int iz = 1; iz < nz - 1; ++iz) {
            const size_t i = (size_t)ix * nz + iz;
            out[i] = idx2 * (image[i + nz] - 2.0f * image[i] + image[i - nz])
                   + idz2 * (image[i + 1] - 2.0f * image[i] + image[i - 1]);
        }
    memcpy(image, out, n * sizeof(float));
    free(out);
}
