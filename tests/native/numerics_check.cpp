// Prints the engine library's CFL bound and points per wavelength
// (rtm::max_stable_dt, rtm::points_per_wavelength; include/rtm/rtm_types.hpp)
// for the cases on the command line, "vmax:vmin:f0:dx:dz:order", one line
// each with the two floats as hexadecimal bit patterns, so the test can
// compare them with the oracle bit for bit.  Host code only: no device.
#include <cstdio>
#include <cstring>

#include "rtm/rtm_types.hpp"

int main(int argc, char** argv) {
    for (int i = 1; i < argc; ++i) {
        float vmax = 0, vmin = 0, f0 = 0, dx = 0, dz = 0;
        int order = 0;
        if (std::sscanf(argv[i], "%f:%f:%f:%f:%f:%d", &vmax, &vmin, &f0, &dx, &dz, &order) != 6) return 2;
        rtm::Grid g;
        g.dx = dx;
        g.dz = dz;
        g.half = order / 2;
        const float dt = rtm::max_stable_dt(vmax, g, order);
        const float ppw = rtm::points_per_wavelength(vmin, f0, g);
        unsigned a, b;
        std::memcpy(&a, &dt, 4);
        std::memcpy(&b, &ppw, 4);
        std::printf("%s %08x %08x\n", argv[i], a, b);
    }
    return 0;
}
