/*
 * rtm_oracle_internal.h -- shared state of the CPU oracle (test infrastructure).
 * Not part of the product; see rtm_oracle.h.
 */
#ifndef RTM_ORACLE_INTERNAL_H
#define RTM_ORACLE_INTERNAL_H

#include "rtm_oracle.h"
#include <stddef.h>

#define ORC_MAX_HALF 6

struct orc_engine {
    int precision;          /* 32 or 64 */
    int variant;            /* 0 cpu, 1 cpu-opt */
    /* grid (rtm::Grid, include/rtm_types.hpp:16-29) */
    int   nx, nz, nb, half, nxe, nze;
    float dx, dz, ox, oz;
    size_t n_ext, n_int;
    /* params (rtm::RTMParams) and time axis */
    int   order, store, verbose;
    float alpha, f0;
    int   nt;
    float dt;
    int   nsnap;
    /* tables + three wavefields, element type float or double per precision */
    void* vdt2;
    void* sponge;
    void* wavelet;
    void* buf[3];
    float  c0_f, cx_f[ORC_MAX_HALF + 1], cz_f[ORC_MAX_HALF + 1];
    double c0_d, cx_d[ORC_MAX_HALF + 1], cz_d[ORC_MAX_HALF + 1];
    /* geometry of the last mapped shot */
    size_t  src;
    size_t* rec;
    int     nrec, nrec_cap;
    /* stage wall clocks (rtm::StageTimes) */
    double t_fwd, t_bwd, t_img;
};

void   orc_set_error(const char* msg);
double orc_now(void);
int    orc_map_geometry_internal(orc_engine* e, const orc_shot* s);

/* cpu-opt pieces (rtm_oracle_opt.c) */
int    orc_opt_num_threads(void);
void   orc_opt_fused_step(const orc_engine* e, const float* p_prev, const float* p_cur,
                          float* p_next);
void   orc_opt_save_snapshot(const orc_engine* e, const float* p_ext, float* dst);
void   orc_opt_imaging(const orc_engine* e, const float* fwd, const float* bwd_ext,
                       float* image, float* illum);

#endif
