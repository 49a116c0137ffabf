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

// Kernel families of the fused step.
//   kLdsTile: a 256-thread workgroup stages a 64 x tile_x tile of p^n (+halo)
//             in LDS; tiles start at cell (0, 0) and cover the padded grid.
//   kMarch:   each wave64 owns a depth strip of 64 cells whose 64 - 2H middle
//             lanes are outputs (strips overlap by 2H), and tile_x columns;
//             it loads every operand of its tile up front (no LDS, no
//             barrier) and takes depth neighbours from other lanes through
//             ds_bpermute (kMarch) or chained DPP row shifts (kMarchDpp).
enum class Family : int { kLdsTile = 0, kMarch = 1, kMarchDpp = 2 };

// Tile grid of one launch.  Tiles are numbered x-major over (ntz x ntx):
// tile = tx * ntz + tz; the injection table is indexed the same way.
struct StepGeometry {
    Family family = Family::kMarch;
    int tile_x = 16;     // output columns per tile
    int tile_z = 64;     // output depth cells per tile
    int x_origin = 0;    // first column of tile 0
    int z_origin = 0;    // first depth cell of tile 0
    int ntz = 0, ntx = 0, ntiles = 0;
    int blocks = 0;      // workgroups launched
};

StepGeometry step_geometry(Family family, int half, int tile_x, int nxe, int nze, int ldz);

inline int tile_of_cell(const StepGeometry& g, int ixe, int ize) {
    return ((ixe - g.x_origin) / g.tile_x) * g.ntz + (ize - g.z_origin) / g.tile_z;
}

// Launch one fused step on `stream` with the geometry from step_geometry.
// With start/stop events, they record the kernel's own start and end.
hipError_t launch_step(Dir dir, int half, const StepGeometry& g, const StepArgs& a, hipStream_t stream,
                       hipEvent_t start = nullptr, hipEvent_t stop = nullptr);

}  // namespace rtm_hip
