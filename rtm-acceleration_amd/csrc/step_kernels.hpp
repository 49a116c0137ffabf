// step_kernels.hpp -- launch interface of the fused leapfrog step kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstddef>

namespace rtm_hip {

constexpr int kMaxHalf = 6;
constexpr int kTileZ = 64;        // one wave64 spans 64 consecutive depth cells
constexpr int kWaves = 4;         // waves per workgroup, each owns tile_x/4 columns
constexpr int kBlockThreads = kTileZ * kWaves;

// Stencil coefficients pre-divided by the spacing (src/rtm/rtm_cpu.cpp:75-79).
struct Coeffs {
    float c0;
    float cx[kMaxHalf + 1];
    float cz[kMaxHalf + 1];
};

// An injection/recording entry of the per-tile table: the extended-grid cell
// (device index ix*ldz+iz) and a code: -1 = add vdt2*wavelet[it] (source),
// >= 0 = receiver ir (forward: record the cell into rec_out[ir*nt+it];
// backward: add vdt2*traces[ir*trace_stride+it]).
struct Entry {
    int cell;
    int code;
};

enum class Dir : int { Forward = 0, Backward = 1 };

struct StepArgs {
    // fields: p_prev holds p^{n-1} damped once; p_cur p^n; p_next p^{n+1}
    const float* p_prev;
    const float* p_cur;
    float*       p_next;
    const float* vdt2;        // [nxe][ldz]
    const float* wx;          // [nxe]  Cerjan taper along x
    const float* wz;          // [ldz]  Cerjan taper along z (0 in the pitch pad)
    Coeffs       coef;
    int nxe, nze, ldz, nb, nx, nz;
    int ntz, ntx, ntiles;     // tile grid: ntz = ldz/64 rows of tiles along z
    int it;
    // per-tile injection / recording table (CSR over logical tiles)
    const int*   tile_ptr;    // [ntiles+1]
    const Entry* entries;
    const float* wavelet;     // [nt]   forward source amplitude
    const float* traces;      // backward: trace-major receiver data
    long long    trace_stride;
    float*       rec_out;     // forward modelling: trace-major, stride rec_stride
    long long    rec_stride;
    // forward: store the new field's interior into snap_out (nx*nz), or NULL
    float*       snap_out;
    // backward: image the new field against snap_in (nx*nz), or NULL
    const float* snap_in;
    float*       image;
    float*       illum;
};

// Launch one fused step on `stream`; tile_x in {4, 8, 16}.
hipError_t launch_step(Dir dir, int half, int tile_x, const StepArgs& a, hipStream_t stream);

// Number of workgroups of a step launch (one per 64 x tile_x tile).
inline int step_tiles_z(int ldz) { return ldz / kTileZ; }
inline int step_tiles_x(int nxe, int tile_x) { return (nxe + tile_x - 1) / tile_x; }

}  // namespace rtm_hip
