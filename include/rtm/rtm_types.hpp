// rtm_types.hpp -- geometry, parameters and shot records of the RTM engine.
//
// Same public surface as the reference's include/rtm_types.hpp:1-66 so that
// code written against it compiles unchanged: Grid, TimeAxis, RTMParams,
// VelocityModel, ShotRecord and the shared numerics.
//
// Layout: every array is depth-fastest.  The interior (physical) grid is
// nx*nz, index ix*nz+iz; the extended grid adds nb absorbing cells on each
// side, nxe*nze, index ix*nze+iz.  Traces are trace-major, traces[ir*nt+it].
#pragma once
#include <cstddef>
#include <string>
#include <vector>

namespace rtm {

constexpr int MAX_HALF = 6;   // spatial orders 2..12

struct Grid {
    int   nx = 0, nz = 0;          // interior samples
    float dx = 10.0f, dz = 10.0f;  // spacing, metres
    float ox = 0.0f, oz = 0.0f;    // origin of the interior, metres
    int   nb = 60;                 // absorbing pad, cells
    int   half = 4;                // stencil half width = order/2

    int nxe() const { return nx + 2 * nb; }
    int nze() const { return nz + 2 * nb; }
    std::size_t n_interior() const { return static_cast<std::size_t>(nx) * nz; }
    std::size_t n_extended() const { return static_cast<std::size_t>(nxe()) * nze(); }
    std::size_t eidx(int ixe, int ize) const { return static_cast<std::size_t>(ixe) * nze() + ize; }
    std::size_t iidx(int ix, int iz) const { return static_cast<std::size_t>(ix) * nz + iz; }
};

struct TimeAxis {
    int   nt = 0;
    float dt = 0.0f;
};

struct RTMParams {
    int   nb             = 60;       // sponge width, cells
    int   order          = 8;        // spatial FD order (2..12, even)
    float sponge_alpha   = 0.0053f;  // Cerjan coefficient
    int   store_interval = 1;        // keep the source wavefield every N steps
    float f0             = 12.0f;    // Ricker peak frequency, Hz
    bool  verbose        = true;
};

struct VelocityModel {
    Grid grid;
    std::vector<float> v;   // interior, m/s, v[ix*nz+iz]
    float vmin() const;
    float vmax() const;
};

struct ShotRecord {
    float sx = 0.0f, sz = 0.0f;     // source, metres
    std::vector<float> rx, rz;      // receivers, metres
    int   nt = 0;
    float dt = 0.0f;
    std::vector<float> traces;      // nrec*nt, trace-major
    int nrec() const { return static_cast<int>(rx.size()); }
    void allocate() { traces.assign(static_cast<std::size_t>(nrec()) * nt, 0.0f); }
};

// ---- shared numerics ---------------------------------------------------------
const float* stencil_coefficients(int order);   // half+1 values; throws on a bad order
int   stencil_half(int order);
float stencil_abs_sum(int order);               // |c0| + 2 sum |ck|
float ricker(float t, float f0);                // t relative to the peak
float max_stable_dt(float vmax, const Grid& g, int order);
float points_per_wavelength(float vmin, float f0, const Grid& g);

}  // namespace rtm
