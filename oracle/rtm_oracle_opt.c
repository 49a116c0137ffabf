/*
 * rtm_oracle_opt.c -- restatement of the reference's "cpu-opt" engine
 * (src/cpu_opt/cpu_opt.cpp:147-312), FP32 only.
 *
 * TEST INFRASTRUCTURE ONLY (see rtm_oracle.h); also the CPU baseline that
 * bench.py times.  Built -O3 -fopenmp -ffp-contract=off (the reference uses
 * -O3 -march=native -ffp-contract=off, CMakeLists.txt:76-78; this build
 * targets x86-64-v3 instead of "native" so the library runs on any host that
 * loads it).  Bit-identical to the "cpu" variant:
 *   - the sponge multiply of p_next is applied as it is written;
 *   - the reference's in-place damping of p_cur is deferred to the next
 *     step's read of p_prev (p_prev*w), the same single multiply;
 *   - injection cells have w == 1 exactly, so injecting after the damped
 *     write equals damping after injection (cpu_opt.cpp:188-189).
 */
#include "rtm_oracle_internal.h"

#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

int orc_opt_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Fused stencil + deferred damping, one instance per half width so the k
 * loop unrolls and the iz loop vectorises (cpu_opt.cpp:147-172). */
#define ORC_FUSED(H)                                                                \
    static void fused_h##H(int nxe, int nze, float c0, const float* cx,             \
                           const float* cz, const float* restrict pp,              \
                           const float* restrict pc, float* restrict pn,           \
                           const float* restrict vdt2, const float* restrict w) {  \
        _Pragma("omp parallel for schedule(static)")                                \
        for (int ix = H; ix < nxe - H; ++ix) {                                      \
            const size_t row = (size_t)ix * nze;                                    \
            for (int iz = H; iz < nze - H; ++iz) {                                  \
                const size_t i = row + iz;                                          \
                float lap = c0 * pc[i];                                             \
                for (int k = 1; k <= H; ++k) {                                      \
                    const size_t kx = (size_t)k * nze;                              \
                    const float tx = cx[k] * (pc[i + kx] + pc[i - kx]);             \
                    const float tz = cz[k] * (pc[i + k] + pc[i - k]);               \
                    lap = lap + (tx + tz);                                          \
                }                                                                   \
                const float prev_damped = pp[i] * w[i];                             \
                const float val = (2.0f * pc[i] - prev_damped) + vdt2[i] * lap;     \
                pn[i] = val * w[i];                                                 \
            }                                                                       \
        }                                                                           \
    }

ORC_FUSED(1)
ORC_FUSED(2)
ORC_FUSED(3)
ORC_FUSED(4)
ORC_FUSED(5)
ORC_FUSED(6)

void orc_opt_fused_step(const orc_engine* e, const float* pp, const float* pc, float* pn) {
    const float* vdt2 = (const float*)e->vdt2;
    const float* w = (const float*)e->sponge;
    switch (e->half) {
        case 1: fused_h1(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        case 2: fused_h2(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        case 3: fused_h3(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        case 4: fused_h4(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        case 5: fused_h5(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        case 6: fused_h6(e->nxe, e->nze, e->c0_f, e->cx_f, e->cz_f, pp, pc, pn, vdt2, w); break;
        default: break;
    }
}

/* Column-parallel interior copy (cpu_opt.cpp:286-293). */
void orc_opt_save_snapshot(const orc_engine* e, const float* p_ext, float* dst) {
    const int nb = e->nb, nz = e->nz, nze = e->nze, nx = e->nx;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ++ix)
        memcpy(dst + (size_t)ix * nz, p_ext + (size_t)(ix + nb) * nze + nb,
               (size_t)nz * sizeof(float));
}

/* Column-parallel imaging (cpu_opt.cpp:370-389). */
void orc_opt_imaging(const orc_engine* e, const float* fwd, const float* bwd_ext,
                     float* image, float* illum) {
    const int nx = e->nx, nz = e->nz, nb = e->nb, nze = e->nze;
#pragma omp parallel for schedule(static)
    for (int ix = 0; ix < nx; ++ix) {
        const float* restrict f = fwd + (size_t)ix * nz;
        const float* restrict b = bwd_ext + (size_t)(ix + nb) * nze + nb;
        float* restrict I = image + (size_t)ix * nz;
        float* restrict S = illum + (size_t)ix * nz;
        for (int iz = 0; iz < nz; ++iz) {
            const float fi = f[iz];
            I[iz] = I[iz] + fi * b[iz];
            S[iz] = S[iz] + fi * fi;
        }
    }
}
