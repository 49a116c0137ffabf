// rtm -- 2D acoustic reverse time migration on one MI355X.
//
// Same flags, defaults, sidecar precedence, checks and outputs as the
// reference's `rtm` (src/main.cpp:18-257); engines come from make_engine
// ("hip", "hip-fp16", "hip-bf16").
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtm/benchmark.hpp"
#include "rtm/comparison.hpp"
#include "rtm/rtm_factory.hpp"
#include "rtm/rtm_hip_engine.hpp"
#include "rtm_hip.h"
#include "rtm/rtm_io.hpp"

using namespace rtm;

namespace {

void usage() {
    std::printf(
        "rtm -- 2D acoustic Reverse Time Migration (MI355X)\n\n"
        "Engine:\n"
        "  --engine NAME          hip | hip-fp16 | hip-bf16      (default hip)\n"
        "  --list-engines         print the engines in this build and exit\n"
        "  --gpus N               GPUs (this build: 1)\n\n"
        "Required:\n"
        "  --velocity PATH        float32 velocity model (m/s)\n"
        "  --shots PATH           shot gathers (.bin = RTMS raw, .sgy/.segy = SEG-Y)\n"
        "  --output PATH          migrated image, float32, image[ix*nz+iz]\n\n"
        "Velocity geometry (read from PATH.hdr when present; CLI overrides):\n"
        "  --nx N  --nz N  --dx M  --dz M  --ox M  --oz M  --vel-layout zfast|xfast\n"
        "  --vel-header PATH      explicit sidecar\n\n"
        "Modelling / migration parameters:\n"
        "  --order N (8)  --nb N (60)  --sponge-alpha F (0.0053)  --f0 F (12)\n"
        "  --store-interval N (1)  --nt N  --max-shots N  --mute-direct V (0 = off)\n\n"
        "Outputs:\n"
        "  --illumination  --filter laplacian  --write-illum-map\n"
        "  --benchmark PATH  --benchmark-csv PATH  --dataset NAME  --quiet\n");
}

std::string file_stem(const std::string& path) {
    const auto slash = path.find_last_of('/');
    std::string name = slash == std::string::npos ? path : path.substr(slash + 1);
    const auto dot = name.find_last_of('.');
    return dot == std::string::npos ? name : name.substr(0, dot);
}

}  // namespace

int main(int argc, char** argv) try {
    std::string vel_path, shots_path, out_path, bench_path, hdr_path, filter = "none";
    std::string engine_name = "hip", bench_csv_path, dataset_label;
    int gpus = 1, nt_override = 0, max_shots = 0;
    io::VelocityHeader vh;
    RTMParams par;
    bool set_nx = false, set_nz = false, set_dx = false, set_dz = false, set_ox = false, set_oz = false;
    bool set_layout = false, want_illum = false, want_illum_map = false;
    float mute_v = 0.0f, segy_sz = 10.0f, segy_rz = 10.0f;

    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
            return argv[++i];
        };
        if (a == "--velocity") vel_path = val();
        else if (a == "--shots") shots_path = val();
        else if (a == "--output") out_path = val();
        else if (a == "--vel-header") hdr_path = val();
        else if (a == "--nx") { vh.nx = std::stoi(val()); set_nx = true; }
        else if (a == "--nz") { vh.nz = std::stoi(val()); set_nz = true; }
        else if (a == "--dx") { vh.dx = std::stof(val()); set_dx = true; }
        else if (a == "--dz") { vh.dz = std::stof(val()); set_dz = true; }
        else if (a == "--ox") { vh.ox = std::stof(val()); set_ox = true; }
        else if (a == "--oz") { vh.oz = std::stof(val()); set_oz = true; }
        else if (a == "--vel-layout") { vh.layout = val(); set_layout = true; }
        else if (a == "--order") par.order = std::stoi(val());
        else if (a == "--nb") par.nb = std::stoi(val());
        else if (a == "--sponge-alpha") par.sponge_alpha = std::stof(val());
        else if (a == "--f0") par.f0 = std::stof(val());
        else if (a == "--store-interval") par.store_interval = std::stoi(val());
        else if (a == "--nt") nt_override = std::stoi(val());
        else if (a == "--max-shots") max_shots = std::stoi(val());
        else if (a == "--mute-direct") mute_v = std::stof(val());
        else if (a == "--illumination") want_illum = true;
        else if (a == "--write-illum-map") want_illum_map = true;
        else if (a == "--filter") filter = val();
        else if (a == "--benchmark") bench_path = val();
        else if (a == "--benchmark-csv") bench_csv_path = val();
        else if (a == "--dataset") dataset_label = val();
        else if (a == "--segy-src-depth") segy_sz = std::stof(val());
        else if (a == "--segy-rec-depth") segy_rz = std::stof(val());
        else if (a == "--quiet") par.verbose = false;
        else if (a == "--engine") engine_name = val();
        else if (a == "--gpus") gpus = std::stoi(val());
        else if (a == "--help" || a == "-h") { usage(); return 0; }
        else if (a == "--list-engines") {
            for (const auto& n : list_engines()) std::printf("%s\n", n.c_str());
            return 0;
        } else throw std::runtime_error("unknown option: " + a);
    }
    if (vel_path.empty() || shots_path.empty() || out_path.empty()) {
        usage();
        throw std::runtime_error("--velocity, --shots and --output are required");
    }

    // ---- data loading: CLI flag > --vel-header > <velocity>.hdr -----------
    auto engine = make_engine(engine_name, gpus);
    Timer io_timer;
    {
        io::VelocityHeader side;
        bool found = false;
        if (!hdr_path.empty()) {
            if (!io::read_header_file(hdr_path, side))
                throw std::runtime_error("cannot read velocity header: " + hdr_path);
            found = true;
        } else {
            found = io::read_header_file(vel_path + ".hdr", side);
        }
        if (found) {
            if (!set_nx) vh.nx = side.nx;
            if (!set_nz) vh.nz = side.nz;
            if (!set_dx) vh.dx = side.dx;
            if (!set_dz) vh.dz = side.dz;
            if (!set_ox) vh.ox = side.ox;
            if (!set_oz) vh.oz = side.oz;
            if (!set_layout) vh.layout = side.layout;
        }
    }
    VelocityModel model = io::read_velocity_raw(vel_path, vh);
    std::vector<ShotRecord> shots = io::read_shots(shots_path, segy_sz, segy_rz);
    if (max_shots > 0 && (int)shots.size() > max_shots) shots.resize(max_shots);
    if (mute_v > 0.0f)
        for (auto& s : shots) mute_direct_wave(s, mute_v, 1.0f / par.f0, 1.0f / par.f0);
    engine->times.io = io_timer.elapsed();

    TimeAxis ta;
    ta.nt = nt_override > 0 ? std::min(nt_override, shots[0].nt) : shots[0].nt;
    ta.dt = shots[0].dt;

    // ---- setup + checks ------------------------------------------------------
    engine->setup(model, par, ta);
    const Grid& g = engine->grid();
    const float dt_max = max_stable_dt(model.vmax(), g, par.order);
    const float ppw = points_per_wavelength(model.vmin(), par.f0, g);
    std::printf("\n=== MODEL ===\n"
                "  nx = %d, nz = %d, dx = %.2f m, dz = %.2f m, origin = (%.1f, %.1f)\n"
                "  vmin = %.1f m/s, vmax = %.1f m/s\n"
                "  nt = %d, dt = %.6f s, T = %.3f s\n"
                "  order = %d, nb = %d, f0 = %.2f Hz, store_interval = %d\n"
                "  shots = %zu, receivers/shot = %d\n"
                "  CFL: dt = %.6f s, dt_max = %.6f s (ratio %.3f) %s\n"
                "  Dispersion: %.2f points per shortest wavelength %s\n"
                "  Snapshot memory: %.1f MiB\n\n",
                g.nx, g.nz, g.dx, g.dz, g.ox, g.oz, model.vmin(), model.vmax(), ta.nt, ta.dt, ta.nt * ta.dt,
                par.order, par.nb, par.f0, par.store_interval, shots.size(), shots[0].nrec(), ta.dt, dt_max,
                ta.dt / dt_max, ta.dt <= dt_max ? "OK" : "*** UNSTABLE ***", ppw,
                ppw >= 4.0f ? "OK" : "*** DISPERSIVE ***", engine->snapshot_bytes() / (1024.0 * 1024.0));
    if (ta.dt > dt_max)
        throw std::runtime_error("time step violates the CFL condition -- reduce dt, reduce the FD order, "
                                 "or coarsen f0/dx");
    if (ppw < 3.0f)
        std::fprintf(stderr, "WARNING: fewer than 3 points per wavelength; the image will be visibly "
                             "dispersive.\n");

    // ---- migration -------------------------------------------------------
    std::vector<float> image, illum;
    std::printf("=== MIGRATION (%s) ===\n", engine->name());
    engine->migrate(shots, image, illum);

    // ---- outputs: the raw zero-lag cross-correlation is the reference image
    io::write_raw_floats(out_path, image);
    std::printf("\nwrote %s  (%d x %d float32, image[ix*nz+iz])\n", out_path.c_str(), g.nx, g.nz);
    const auto dot = out_path.find_last_of('.');
    const std::string stem = dot == std::string::npos ? out_path : out_path.substr(0, dot);
    if (want_illum_map) io::write_raw_floats(stem + "_illummap.bin", illum);
    // post-processed copies (src/main.cpp:211-222): on the device, from the
    // image the HIP engine keeps resident after migrate
    auto* hip = dynamic_cast<HipRTM*>(engine.get());
    if (want_illum) {
        std::vector<float> tmp;
        if (hip) {
            hip->postprocess(RTM_HIP_POST_ILLUM, 1e-4f, tmp);
        } else {
            tmp = image;
            illumination_compensation(tmp, illum, 1e-4f);
        }
        io::write_raw_floats(stem + "_illum.bin", tmp);
        std::printf("wrote %s_illum.bin\n", stem.c_str());
    }
    if (filter == "laplacian") {
        std::vector<float> tmp;
        if (hip) {
            hip->postprocess(RTM_HIP_POST_LAPLACIAN | (want_illum ? RTM_HIP_POST_ILLUM : 0), 1e-4f, tmp);
        } else {
            tmp = image;
            if (want_illum) illumination_compensation(tmp, illum, 1e-4f);
            laplacian_filter(tmp, g);
        }
        io::write_raw_floats(stem + "_filtered.bin", tmp);
        std::printf("wrote %s_filtered.bin\n", stem.c_str());
    }

    // ---- benchmark report -------------------------------------------------
    BenchmarkContext ctx;
    ctx.nx = g.nx; ctx.nz = g.nz; ctx.nb = g.nb; ctx.dx = g.dx; ctx.dz = g.dz;
    ctx.nt = ta.nt; ctx.dt = ta.dt; ctx.nshots = (int)shots.size(); ctx.order = par.order;
    ctx.store_interval = par.store_interval;
    ctx.snapshot_bytes = engine->snapshot_bytes();
    ctx.results.push_back({engine->name(), engine->times, engine->is_gpu()});
    print_benchmark_report(std::cout, ctx);
    if (!bench_path.empty()) {
        std::ofstream bf(bench_path, std::ios::app);
        print_benchmark_report(bf, ctx);
    }
    if (!bench_csv_path.empty()) {
        RunInfo info;
        info.host = host_name();
        info.cpu_name = cpu_model_name();
        info.gpu_name = engine->device_name();
        info.engine = engine_name;
        info.dataset = dataset_label.empty() ? file_stem(shots_path) : dataset_label;
        info.threads = engine->num_threads();
        info.gpus = engine->num_devices();
        append_benchmark_csv(bench_csv_path, ctx, ctx.results.back(), info);
        std::printf("appended benchmark row to %s\n", bench_csv_path.c_str());
    }
    return 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "\nERROR: %s\n", e.what());
    return 1;
}
