// engine_base.cpp -- default drivers of RTMEngine and image post-processing.
//
// migrate: zero image, illumination and the host snapshot store, then for
// every shot in order forward (filling the store) and backward (imaging
// against it) -- a running FP32 sum over shots (src/rtm/rtm.cpp:83-103).
// model_shot: forward propagation with recording (src/rtm/rtm.cpp:105-111).
#include <algorithm>
#include <cstdio>

#include "rtm/benchmark.hpp"
#include "rtm/rtm_engine.hpp"

namespace rtm {

void RTMEngine::migrate(const std::vector<ShotRecord>& shots, std::vector<float>& image,
                        std::vector<float>& illumination) {
    Timer total;
    const std::size_t n = g_.n_interior();
    image.assign(n, 0.0f);
    illumination.assign(n, 0.0f);
    snap_.assign(static_cast<std::size_t>(nsnap_) * n, 0.0f);
    for (std::size_t is = 0; is < shots.size(); ++is) {
        if (par_.verbose) {
            std::printf("  shot %2zu / %2zu   sx=%8.1f m  sz=%6.1f m  nrec=%d\n", is + 1, shots.size(),
                        shots[is].sx, shots[is].sz, shots[is].nrec());
            std::fflush(stdout);
        }
        forward_propagation(shots[is], &snap_, nullptr);
        backward_propagation(shots[is], snap_, image, illumination);
    }
    times.total += total.elapsed();
}

void RTMEngine::model_shot(ShotRecord& shot) {
    shot.nt = ta_.nt;
    shot.dt = ta_.dt;
    std::vector<float> rec(static_cast<std::size_t>(shot.nrec()) * ta_.nt, 0.0f);
    forward_propagation(shot, nullptr, &rec);
    shot.traces.swap(rec);
}

// I / (S + eps*max S + 1e-30)  (src/rtm/imaging.cpp:48-56).
void illumination_compensation(std::vector<float>& image, const std::vector<float>& illumination,
                               float eps_relative) {
    float smax = 0.0f;
    for (float s : illumination) smax = std::max(smax, s);
    const float eps = eps_relative * smax + 1e-30f;
    for (std::size_t i = 0; i < image.size(); ++i) image[i] /= (illumination[i] + eps);
}

// Second-order Laplacian high-pass of the image; edges become 0
// (src/rtm/imaging.cpp:63-75).
void laplacian_filter(std::vector<float>& image, const Grid& g) {
    std::vector<float> out(image.size(), 0.0f);
    const float idx2 = 1.0f / (g.dx * g.dx);
    const float idz2 = 1.0f / (g.dz * g.dz);
    for (int ix = 1; ix < g.nx - 1; ++ix) {
        const float* c = image.data() + static_cast<std::size_t>(ix) * g.nz;
        float* o = out.data() + static_cast<std::size_t>(ix) * g.nz;
        for (int iz = 1; iz < g.nz - 1; ++iz) {
            o[iz] = idx2 * (c[iz + g.nz] - 2.0f * c[iz] + c[iz - g.nz]) +
                    idz2 * (c[iz + 1] - 2.0f * c[iz] + c[iz - 1]);
        }
    }
    image.swap(out);
}

}  // namespace rtm
