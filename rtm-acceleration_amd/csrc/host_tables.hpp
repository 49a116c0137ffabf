// host_tables.hpp -- host-side tables uploaded by the device engine.
#pragma once
#include <vector>
#include "step_kernels.hpp"

namespace rtm_hip {

struct HostTables {
    int nxe = 0, nze = 0, ldz = 0;
    std::vector<float> vdt2;     // [nxe + kGuardCols][ldz], (v dt)^2, 0 in the pitch pad and guard
    std::vector<float> wx;       // [nxe + kGuardCols], 0 in the guard
    std::vector<float> wz;       // [ldz], 0 in the pitch pad
    std::vector<float> wavelet;  // [nt]
    Coeffs coef{};
};

// Builds the tables of CPUReferenceRTM::setup (src/rtm/rtm_cpu.cpp:33-79) with
// its float formulas; the 2-D sponge is kept separable (the reference's table
// is the float product wx[ix]*wz[iz], which the kernel forms on the fly).
HostTables build_tables(int nx, int nz, float dx, float dz, int nb, int order, float alpha,
                        float f0, int nt, float dt, const float* v, int ldz);

}  // namespace rtm_hip
