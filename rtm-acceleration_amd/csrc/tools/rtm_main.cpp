// rtm -- 2D acoustic reverse time migration on one MI355X.
//
// The command line of the reference's `rtm` (flags and defaults of
// src/main.cpp:18-52): the velocity geometry comes from a flag, else from
// --vel-header, else from <velocity>.hdr; a time step beyond the CFL limit is
// refused and fewer than 3 points per wavelength draw a warning; the outputs
// are the raw zero-lag cross-correlation image and, on request,
// <stem>_illummap.bin, <stem>_illum.bin, <stem>_filtered.bin, the benchmark
// report and one row of the benchmark CSV.  Exit status 0, or 1 with a
// message on any error.  Engines come from make_engine ("hip", "hip-fp16",
// "hip-bf16"); the post-processed images are computed on the device from the
// image the HIP engine keeps there.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "cli.hpp"
#include "rtm/benchmark.hpp"
#include "rtm/rtm_factory.hpp"
#include "rtm/rtm_hip_engine.hpp"
#include "rtm/rtm_io.hpp"
#include "rtm_hip.h"

using namespace rtm;

namespace {

// Everything the command line asks for.
struct Request {
    std::string velocity, shots, output, header, filter = "none";
    std::string report, csv, dataset, engine = "hip";
    int gpus = 1, nt = 0, max_shots = 0;
    io::VelocityHeader geometry;
    RTMParams params;
    float mute_velocity = 0.0f, segy_src_depth = 10.0f, segy_rec_depth = 10.0f;
    bool illum = false, illum_map = false, quiet = false;
};

void declare(cli::Options& o, Request& r) {
    o.section("Engine")
        .text("--engine", r.engine, "hip | hip-fp16 | hip-bf16 (default hip)")
        .action("--list-engines", [] { for (const auto& n : list_engines()) std::printf("%s\n", n.c_str()); },
                "print the engines of this build and exit")
        .integer("--gpus", r.gpus, "GPUs (this build: 1)");
    o.section("Inputs and output (required)")
        .text("--velocity", r.velocity, "float32 velocity model, m/s")
        .text("--shots", r.shots, "shot gathers: .bin = RTMS, .sgy/.segy = SEG-Y")
        .text("--output", r.output, "migrated image, float32, image[ix*nz+iz]");
    o.section("Velocity geometry (a flag wins over --vel-header, which wins over PATH.hdr)")
        .integer("--nx", r.geometry.nx, "samples along x")
        .integer("--nz", r.geometry.nz, "samples along z")
        .real("--dx", r.geometry.dx, "spacing along x, m")
        .real("--dz", r.geometry.dz, "spacing along z, m")
        .real("--ox", r.geometry.ox, "origin x, m")
        .real("--oz", r.geometry.oz, "origin z, m")
        .text("--vel-layout", r.geometry.layout, "zfast | xfast")
        .text("--vel-header", r.header, "explicit geometry sidecar");
    o.section("Modelling and migration")
        .integer("--order", r.params.order, "spatial FD order 2..12 (default 8)")
        .integer("--nb", r.params.nb, "sponge width, cells (default 60)")
        .real("--sponge-alpha", r.params.sponge_alpha, "Cerjan coefficient (default 0.0053)")
        .real("--f0", r.params.f0, "Ricker peak frequency, Hz (default 12)")
        .integer("--store-interval", r.params.store_interval, "keep the source wavefield every N steps (default 1)")
        .integer("--nt", r.nt, "truncate the time axis")
        .integer("--max-shots", r.max_shots, "migrate only the first N shots")
        .real("--mute-direct", r.mute_velocity, "direct-wave mute velocity, m/s (0 = off)");
    o.section("Outputs")
        .flag("--illumination", r.illum, "also write <output>_illum.bin (illumination compensated)")
        .text("--filter", r.filter, "laplacian: also write <output>_filtered.bin")
        .flag("--write-illum-map", r.illum_map, "also write <output>_illummap.bin")
        .text("--benchmark", r.report, "append the benchmark report to PATH")
        .text("--benchmark-csv", r.csv, "append one benchmark CSV row to PATH")
        .text("--dataset", r.dataset, "dataset label of the CSV row (default: shots file stem)")
        .flag("--quiet", r.quiet, "less chatter");
    o.section("SEG-Y input")
        .real("--segy-src-depth", r.segy_src_depth, "source depth, m (default 10)")
        .real("--segy-rec-depth", r.segy_rec_depth, "receiver depth, m (default 10)");
}

// Fills the geometry fields no flag set from the sidecar: the explicit
// --vel-header (which must exist), else <velocity>.hdr when it is there.
void complete_geometry(Request& r, const cli::Options& opts) {
    io::VelocityHeader side;
    if (!r.header.empty()) {
        if (!io::read_header_file(r.header, side)) throw std::runtime_error("cannot read velocity header: " + r.header);
    } else if (!io::read_header_file(r.velocity + ".hdr", side)) {
        return;
    }
    using H = io::VelocityHeader;
    const struct {
        const char* flag;
        void (*take)(H&, const H&);
    } fields[] = {
        {"--nx", [](H& d, const H& s) { d.nx = s.nx; }},
        {"--nz", [](H& d, const H& s) { d.nz = s.nz; }},
        {"--dx", [](H& d, const H& s) { d.dx = s.dx; }},
        {"--dz", [](H& d, const H& s) { d.dz = s.dz; }},
        {"--ox", [](H& d, const H& s) { d.ox = s.ox; }},
        {"--oz", [](H& d, const H& s) { d.oz = s.oz; }},
        {"--vel-layout", [](H& d, const H& s) { d.layout = s.layout; }},
    };
    for (const auto& f : fields)
        if (!opts.given(f.flag)) f.take(r.geometry, side);
}

// "dir/marmousi_shots.bin" -> "marmousi_shots"
std::string stem_of(const std::string& path) {
    std::string name = path.substr(path.find_last_of('/') + 1);
    return name.substr(0, name.find_last_of('.'));
}

// The output path without its extension: the prefix of the extra images.
std::string output_prefix(const std::string& path) {
    const auto dot = path.find_last_of('.');
    return dot == std::string::npos ? path : path.substr(0, dot);
}

struct Inputs {
    VelocityModel model;
    std::vector<ShotRecord> shots;
    TimeAxis axis;
};

Inputs load_inputs(const Request& r) {
    Inputs in;
    in.model = io::read_velocity_raw(r.velocity, r.geometry);
    in.shots = io::read_shots(r.shots, r.segy_src_depth, r.segy_rec_depth);
    if (in.shots.empty()) throw std::runtime_error("no shots in " + r.shots);
    if (r.max_shots > 0 && (int)in.shots.size() > r.max_shots) in.shots.resize((std::size_t)r.max_shots);
    if (r.mute_velocity > 0.0f) {
        const float window = 1.0f / r.params.f0;      // pad and taper of one source period
        for (ShotRecord& s : in.shots) mute_direct_wave(s, r.mute_velocity, window, window);
    }
    // the first shot sets the time axis for all; --nt can only shorten it
    in.axis.dt = in.shots.front().dt;
    in.axis.nt = r.nt > 0 ? std::min(r.nt, in.shots.front().nt) : in.shots.front().nt;
    return in;
}

// Prints what will run and refuses an unstable time step.
void preflight(const Request& r, const RTMEngine& engine, const Inputs& in) {
    const Grid& g = engine.grid();
    const float dt = in.axis.dt;
    const float dt_max = max_stable_dt(in.model.vmax(), g, r.params.order);
    const float ppw = points_per_wavelength(in.model.vmin(), r.params.f0, g);
    const bool stable = dt <= dt_max;
    std::printf("\nmodel       %d x %d cells at %.2f x %.2f m, origin (%.1f, %.1f) m, v %.1f .. %.1f m/s\n"
                "time        nt %d, dt %.6f s, T %.3f s\n"
                "stencil     order %d, sponge %d cells, Ricker %.2f Hz, snapshot every %d steps (%.1f MiB)\n"
                "shots       %zu x %d receivers\n"
                "stability   dt / dt_max = %.3f (dt_max %.6f s): %s\n"
                "dispersion  %.2f points per shortest wavelength: %s\n\n",
                g.nx, g.nz, g.dx, g.dz, g.ox, g.oz, in.model.vmin(), in.model.vmax(), in.axis.nt, dt,
                in.axis.nt * dt, r.params.order, r.params.nb, r.params.f0, r.params.store_interval,
                engine.snapshot_bytes() / (1024.0 * 1024.0), in.shots.size(), in.shots.front().nrec(), dt / dt_max,
                dt_max, stable ? "stable" : "*** UNSTABLE ***", ppw, ppw >= 4.0f ? "ok" : "*** DISPERSIVE ***");
    if (!stable) {
        char msg[256];
        std::snprintf(msg, sizeof msg,
                      "time step dt = %.6g s exceeds the CFL limit %.6g s of order %d: reduce dt, lower the FD "
                      "order, or coarsen f0/dx",
                      dt, dt_max, r.params.order);
        throw std::runtime_error(msg);
    }
    if (ppw < 3.0f)
        std::fprintf(stderr, "WARNING: fewer than 3 points per wavelength; the image will be visibly dispersive.\n");
}

// The raw image and the extra images asked for.  The compensated and
// filtered copies come from the device-resident image when the engine is the
// HIP one, else from the host functions on the downloaded image.
void write_images(const Request& r, RTMEngine& engine, const std::vector<float>& image,
                  const std::vector<float>& illum) {
    const Grid& g = engine.grid();
    io::write_raw_floats(r.output, image);
    std::printf("\nwrote %s  (%d x %d float32, image[ix*nz+iz])\n", r.output.c_str(), g.nx, g.nz);
    const std::string prefix = output_prefix(r.output);
    if (r.illum_map) io::write_raw_floats(prefix + "_illummap.bin", illum);

    auto* hip = dynamic_cast<HipRTM*>(&engine);
    auto derived = [&](const char* suffix, bool compensate, bool filter) {
        std::vector<float> out;
        if (hip) {
            hip->postprocess((compensate ? RTM_HIP_POST_ILLUM : 0) | (filter ? RTM_HIP_POST_LAPLACIAN : 0), 1e-4f,
                             out);
        } else {
            out = image;
            if (compensate) illumination_compensation(out, illum, 1e-4f);
            if (filter) laplacian_filter(out, g);
        }
        io::write_raw_floats(prefix + suffix, out);
        std::printf("wrote %s%s\n", prefix.c_str(), suffix);
    };
    if (r.illum) derived("_illum.bin", true, false);
    if (r.filter == "laplacian") derived("_filtered.bin", r.illum, true);
}

void report(const Request& r, const RTMEngine& engine, const Inputs& in) {
    const Grid& g = engine.grid();
    BenchmarkContext ctx;
    ctx.nx = g.nx;
    ctx.nz = g.nz;
    ctx.nb = g.nb;
    ctx.dx = g.dx;
    ctx.dz = g.dz;
    ctx.nt = in.axis.nt;
    ctx.dt = in.axis.dt;
    ctx.nshots = (int)in.shots.size();
    ctx.order = r.params.order;
    ctx.store_interval = r.params.store_interval;
    ctx.snapshot_bytes = engine.snapshot_bytes();
    ctx.results.push_back({engine.name(), engine.times, engine.is_gpu()});
    print_benchmark_report(std::cout, ctx);
    if (!r.report.empty()) {
        std::ofstream f(r.report, std::ios::app);
        print_benchmark_report(f, ctx);
    }
    if (!r.csv.empty()) {
        RunInfo info;
        info.host = host_name();
        info.cpu_name = cpu_model_name();
        info.gpu_name = engine.device_name();
        info.engine = r.engine;
        info.dataset = r.dataset.empty() ? stem_of(r.shots) : r.dataset;
        info.threads = engine.num_threads();
        info.gpus = engine.num_devices();
        append_benchmark_csv(r.csv, ctx, ctx.results.back(), info);
        std::printf("appended benchmark row to %s\n", r.csv.c_str());
    }
}

int run(int argc, char** argv) {
    Request r;
    cli::Options opts("rtm -- 2D acoustic reverse time migration on one MI355X\n"
                      "usage: rtm --velocity PATH --shots PATH --output PATH [options]");
    declare(opts, r);
    opts.action("--help", [&opts] { opts.usage(stdout); }, "this text");
    const auto rest = opts.parse(argc, argv);
    if (opts.stopped()) return 0;
    if (!rest.empty()) throw std::runtime_error("unknown option: " + rest.front());
    if (r.velocity.empty() || r.shots.empty() || r.output.empty()) {
        opts.usage(stdout);
        throw std::runtime_error("--velocity, --shots and --output are required");
    }
    r.params.verbose = !r.quiet;

    auto engine = make_engine(r.engine, r.gpus);
    Timer io_clock;
    complete_geometry(r, opts);
    const Inputs in = load_inputs(r);
    engine->times.io = io_clock.elapsed();

    engine->setup(in.model, r.params, in.axis);
    preflight(r, *engine, in);

    std::vector<float> image, illum;
    std::printf("migration   %s\n", engine->name());
    engine->migrate(in.shots, image, illum);
    write_images(r, *engine, image, illum);
    report(r, *engine, in);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "\nERROR: %s\n", e.what());
        return 1;
    }
}
