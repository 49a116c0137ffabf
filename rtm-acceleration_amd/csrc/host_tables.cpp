// host_tables.cpp -- see host_tables.hpp.
#include "host_tables.hpp"

#include <algorithm>
#include <cmath>

#include "rtm/rtm_types.hpp"

namespace rtm_hip {

namespace {
// Cerjan taper: exp(-(alpha (nb-d))^2) within nb cells of the edge, exactly 1
// elsewhere; d is the distance to the nearer edge (src/rtm/rtm_cpu.cpp:50-56).
float cerjan(int i, int n, int nb, float alpha) {
    const int d = std::min(i, n - 1 - i);
    if (d >= nb) return 1.0f;
    const float a = alpha * static_cast<float>(nb - d);
    return std::exp(-a * a);
}
}  // namespace

HostTables build_tables(int nx, int nz, float dx, float dz, int nb, int order, float alpha,
                        float f0, int nt, float dt, const float* v, int ldz) {
    HostTables t;
    t.nxe = nx + 2 * nb;
    t.nze = nz + 2 * nb;
    t.ldz = ldz;

    // (v dt)^2 with edge replication into the pad (src/rtm/rtm_cpu.cpp:33-42)
    const float dt2 = dt * dt;
    t.vdt2.assign(static_cast<std::size_t>(t.nxe + kGuardCols) * ldz, 0.0f);
    for (int ixe = 0; ixe < t.nxe; ++ixe) {
        const int ix = std::clamp(ixe - nb, 0, nx - 1);
        float* col = t.vdt2.data() + static_cast<std::size_t>(ixe) * ldz;
        const float* vcol = v + static_cast<std::size_t>(ix) * nz;
        for (int ize = 0; ize < t.nze; ++ize) {
            const float vv = vcol[std::clamp(ize - nb, 0, nz - 1)];
            col[ize] = vv * vv * dt2;
        }
    }

    t.wx.assign(t.nxe + kGuardCols, 0.0f);
    for (int i = 0; i < t.nxe; ++i) t.wx[i] = cerjan(i, t.nxe, nb, alpha);
    t.wz.assign(ldz, 0.0f);
    for (int i = 0; i < t.nze; ++i) t.wz[i] = cerjan(i, t.nze, nb, alpha);

    // Ricker delayed by 1.2/f0 (src/rtm/rtm_cpu.cpp:67-70)
    t.wavelet.resize(nt);
    const float t0 = 1.2f / f0;
    for (int it = 0; it < nt; ++it) t.wavelet[it] = rtm::ricker(static_cast<float>(it) * dt - t0, f0);

    // coefficients pre-divided by the spacing (src/rtm/rtm_cpu.cpp:75-79)
    const float* c = rtm::stencil_coefficients(order);
    const float idx2 = 1.0f / (dx * dx);
    const float idz2 = 1.0f / (dz * dz);
    t.coef.c0 = c[0] * (idx2 + idz2);
    for (int k = 0; k <= kMaxHalf; ++k) t.coef.cx[k] = t.coef.cz[k] = 0.0f;
    for (int k = 1; k <= order / 2; ++k) {
        t.coef.cx[k] = c[k] * idx2;
        t.coef.cz[k] = c[k] * idz2;
    }
    return t;
}

}  // namespace rtm_hip
