// numerics.cpp -- shared numerics of the engine (the reference's
// src/rtm/rtm.cpp:14-77): FD coefficients, Ricker wavelet, CFL limit and
// dispersion figure, with the reference's float formulas so that every table
// derived from them is bit-identical.
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "rtm/rtm_types.hpp"

namespace rtm {

namespace {
// Central second-derivative coefficients {c0, c1..cH}, Taylor-optimal, each set
// summing to zero (src/rtm/rtm.cpp:14-23).  Written as float quotients so the
// rounded values match the reference's float literals.
const float kOrder2[]  = {-2.0f, 1.0f};
const float kOrder4[]  = {-2.5f, 4.0f / 3.0f, -1.0f / 12.0f};
const float kOrder6[]  = {-49.0f / 18.0f, 1.5f, -0.15f, 1.0f / 90.0f};
const float kOrder8[]  = {-205.0f / 72.0f, 8.0f / 5.0f, -1.0f / 5.0f, 8.0f / 315.0f,
                          -1.0f / 560.0f};
const float kOrder10[] = {-5269.0f / 1800.0f, 5.0f / 3.0f, -5.0f / 21.0f, 5.0f / 126.0f,
                          -5.0f / 1008.0f, 1.0f / 3150.0f};
const float kOrder12[] = {-5369.0f / 1800.0f, 12.0f / 7.0f, -15.0f / 56.0f, 10.0f / 189.0f,
                          -1.0f / 112.0f, 2.0f / 1925.0f, -1.0f / 16632.0f};
}  // namespace

const float* stencil_coefficients(int order) {
    switch (order) {
        case 2: return kOrder2;
        case 4: return kOrder4;
        case 6: return kOrder6;
        case 8: return kOrder8;
        case 10: return kOrder10;
        case 12: return kOrder12;
        default:
            throw std::runtime_error("unsupported spatial order (use 2, 4, 6, 8, 10 or 12)");
    }
}

int stencil_half(int order) {
    (void)stencil_coefficients(order);
    return order / 2;
}

float stencil_abs_sum(int order) {
    const float* c = stencil_coefficients(order);
    float s = std::fabs(c[0]);
    for (int k = 1; k <= order / 2; ++k) s += 2.0f * std::fabs(c[k]);
    return s;
}

// (1 - 2a) exp(-a), a = pi^2 f0^2 t^2, all in float (src/rtm/rtm.cpp:52-55).
float ricker(float t, float f0) {
    const float a = static_cast<float>(M_PI * M_PI) * f0 * f0 * t * t;
    return (1.0f - 2.0f * a) * std::exp(-a);
}

// dt <= 2 / (vmax sqrt(S (1/dx^2 + 1/dz^2)))  (src/rtm/rtm.cpp:59-62).
float max_stable_dt(float vmax, const Grid& g, int order) {
    const float S = stencil_abs_sum(order);
    return 2.0f / (vmax * std::sqrt(S * (1.0f / (g.dx * g.dx) + 1.0f / (g.dz * g.dz))));
}

// Points per shortest wavelength with fmax = 2.5 f0 (src/rtm/rtm.cpp:66-70).
float points_per_wavelength(float vmin, float f0, const Grid& g) {
    const float fmax = 2.5f * f0;
    return (vmin / fmax) / std::max(g.dx, g.dz);
}

float VelocityModel::vmin() const { return v.empty() ? 0.0f : *std::min_element(v.begin(), v.end()); }
float VelocityModel::vmax() const { return v.empty() ? 0.0f : *std::max_element(v.begin(), v.end()); }

}  // namespace rtm
