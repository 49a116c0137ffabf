// rtm_synth -- synthetic shot gathers by forward modelling on the GPU.
//
// The command line of the reference's rtm_synth (flags and defaults of
// src/tools/make_synthetic.cpp:28-69): the model is --velocity PATH (geometry
// from PATH.hdr, flags win) or the built-in one, written to
// <outdir>/velocity.bin + .hdr; one surface line of receivers every --rec-dx
// for every shot; shots spread evenly over the line, or at --shot-x0 +
// i*--shot-dx; each shot modelled with model_shot and all written as one
// RTMS file.  The modelling uses the migration's own propagator (an "inverse
// crime", fit for validating the pipeline, not for judging imaging accuracy).
// Exit status 0, or 1 with a message.  The model and the acquisition geometry
// are data formats the reference's scripts depend on, restated exactly.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "cli.hpp"
#include "rtm/benchmark.hpp"
#include "rtm/rtm_factory.hpp"
#include "rtm/rtm_io.hpp"

using namespace rtm;

namespace {

struct Request {
    std::string outdir = "data/synthetic", velocity, shots_out, engine = "hip";
    int nt = 1400, nshots = 4, order = 8, nb = 60;
    float dt = 0.001f, f0 = 12.0f;
    float src_z = 20.0f, rec_z = 20.0f, rec_dx = 20.0f;
    float shot_x0 = -1.0f, shot_dx = -1.0f;     // shot_dx <= 0: spread evenly
    io::VelocityHeader geometry;                // nx, nz, dx, dz of the grid
    Request() {
        geometry.nx = 201;
        geometry.nz = 151;
        geometry.dx = geometry.dz = 10.0f;
    }
};

void declare(cli::Options& o, Request& r) {
    o.section("Model")
        .text("--velocity", r.velocity, "model through this float32 velocity file (default: the built-in model)")
        .text("--vel-layout", r.geometry.layout, "zfast | xfast")
        .integer("--nx", r.geometry.nx, "samples along x (default 201; --velocity: from PATH.hdr)")
        .integer("--nz", r.geometry.nz, "samples along z (default 151)")
        .real("--dx", r.geometry.dx, "spacing along x, m (default 10)")
        .real("--dz", r.geometry.dz, "spacing along z, m (default 10)");
    o.section("Acquisition and modelling")
        .integer("--nt", r.nt, "time samples (default 1400)")
        .real("--dt", r.dt, "time step, s (default 0.001)")
        .real("--f0", r.f0, "Ricker peak frequency, Hz (default 12)")
        .integer("--nshots", r.nshots, "shots (default 4)")
        .real("--shot-x0", r.shot_x0, "first shot x, m (with --shot-dx)")
        .real("--shot-dx", r.shot_dx, "shot spacing, m (default: spread evenly)")
        .real("--rec-dx", r.rec_dx, "receiver spacing, m (default 20)")
        .real("--src-z", r.src_z, "source depth, m (default 20)")
        .real("--rec-z", r.rec_z, "receiver depth, m (default 20)")
        .integer("--order", r.order, "spatial FD order (default 8)")
        .integer("--nb", r.nb, "sponge width, cells (default 60)")
        .text("--engine", r.engine, "hip | hip-fp16 | hip-bf16 (default hip)");
    o.section("Outputs")
        .text("--outdir", r.outdir, "directory of velocity.bin and shots.bin (default data/synthetic)")
        .text("--shots-out", r.shots_out, "shots file (default <outdir>/shots.bin)");
}

// The built-in model: 1500 m/s; 2000 m/s from z = 400 m; 2500 m/s below an
// interface dipping at 0.15 through z = 900 m at mid-line; a 60 x 60 m box at
// 2800 m/s with its corner at (1200 m, 700 m), a point diffractor.
VelocityModel builtin_model(int nx, int nz, float dx, float dz) {
    VelocityModel m;
    m.grid.nx = nx;
    m.grid.nz = nz;
    m.grid.dx = dx;
    m.grid.dz = dz;
    m.v.resize((std::size_t)nx * nz);
    const float mid = 0.5f * nx * dx;
    for (int ix = 0; ix < nx; ++ix) {
        const float x = ix * dx;
        const float dip = 900.0f + 0.15f * (x - mid);
        const bool in_box_x = x >= 1200.0f && x < 1260.0f;
        for (int iz = 0; iz < nz; ++iz) {
            const float z = iz * dz;
            float v = z >= dip ? 2500.0f : (z >= 400.0f ? 2000.0f : 1500.0f);
            if (in_box_x && z >= 700.0f && z < 760.0f) v = 2800.0f;
            m.v[(std::size_t)ix * nz + iz] = v;
        }
    }
    return m;
}

VelocityModel load_or_build(Request& r, const cli::Options& opts) {
    if (r.velocity.empty()) {
        VelocityModel m = builtin_model(r.geometry.nx, r.geometry.nz, r.geometry.dx, r.geometry.dz);
        io::VelocityHeader h = r.geometry;
        h.ox = h.oz = 0.0f;
        h.layout = "zfast";
        io::write_velocity_raw(r.outdir + "/velocity.bin", m);
        io::write_header_file(r.outdir + "/velocity.bin.hdr", h);
        return m;
    }
    // the sidecar fills what no flag set; the layout only when left zfast
    io::VelocityHeader side;
    if (io::read_header_file(r.velocity + ".hdr", side)) {
        if (!opts.given("--nx")) r.geometry.nx = side.nx;
        if (!opts.given("--nz")) r.geometry.nz = side.nz;
        if (!opts.given("--dx")) r.geometry.dx = side.dx;
        if (!opts.given("--dz")) r.geometry.dz = side.dz;
        if (r.geometry.layout == "zfast") r.geometry.layout = side.layout;
        r.geometry.ox = side.ox;
        r.geometry.oz = side.oz;
    }
    return io::read_velocity_raw(r.velocity, r.geometry);
}

// Every shot records on the whole line: nrec = floor((nx-1) dx / rec_dx) + 1
// receivers at ir * rec_dx; shot i at shot_x0 + i shot_dx, or at
// xmax (i+1) / (nshots+1) when no spacing is given.
std::vector<ShotRecord> surface_line(const Request& r, const Grid& g) {
    const float xmax = (g.nx - 1) * g.dx;
    const int nrec = (int)((g.nx - 1) * g.dx / r.rec_dx) + 1;
    std::vector<ShotRecord> shots((std::size_t)r.nshots);
    for (int i = 0; i < r.nshots; ++i) {
        ShotRecord& s = shots[(std::size_t)i];
        s.sx = r.shot_dx > 0.0f ? (r.shot_x0 >= 0.0f ? r.shot_x0 : 0.0f) + r.shot_dx * (float)i
                                : xmax * (float)(i + 1) / (float)(r.nshots + 1);
        s.sz = r.src_z;
        s.nt = r.nt;
        s.dt = r.dt;
        s.rz.assign((std::size_t)nrec, r.rec_z);
        s.rx.resize((std::size_t)nrec);
        for (int k = 0; k < nrec; ++k) s.rx[(std::size_t)k] = k * r.rec_dx;
    }
    return shots;
}

int run(int argc, char** argv) {
    Request r;
    cli::Options opts("rtm_synth -- synthetic shot gathers modelled on one MI355X\n"
                      "usage: rtm_synth [options]");
    declare(opts, r);
    opts.action("--help", [&opts] { opts.usage(stdout); }, "this text");
    const auto rest = opts.parse(argc, argv);
    if (opts.stopped()) return 0;
    if (!rest.empty()) throw std::runtime_error("unknown option: " + rest.front());

    const VelocityModel model = load_or_build(r, opts);
    RTMParams par;
    par.nb = r.nb;
    par.order = r.order;
    par.f0 = r.f0;
    par.store_interval = 1;
    TimeAxis ta;
    ta.nt = r.nt;
    ta.dt = r.dt;
    auto engine = make_engine(r.engine);
    engine->setup(model, par, ta);
    const Grid& g = engine->grid();
    std::vector<ShotRecord> shots = surface_line(r, g);

    const float dt_max = max_stable_dt(model.vmax(), g, r.order);
    const float ppw = points_per_wavelength(model.vmin(), r.f0, g);
    std::printf("modelled shot gathers (the migration's own propagator)\n"
                "  grid %d x %d at %.1f x %.1f m, v %.0f .. %.0f m/s\n"
                "  nt %d, dt %.6f s (stable up to %.6f s), T %.2f s, %.2f points per wavelength\n"
                "  %d shots x %d receivers\n\n",
                g.nx, g.nz, g.dx, g.dz, model.vmin(), model.vmax(), r.nt, r.dt, dt_max, r.nt * r.dt, ppw, r.nshots,
                shots.empty() ? 0 : shots.front().nrec());
    if (r.dt > dt_max) throw std::runtime_error("chosen dt is unstable (CFL)");

    Timer clock;
    for (int i = 0; i < r.nshots; ++i) {
        std::printf("  shot %d / %d at x = %.1f m\n", i + 1, r.nshots, shots[(std::size_t)i].sx);
        std::fflush(stdout);
        engine->model_shot(shots[(std::size_t)i]);
    }
    const std::string path = r.shots_out.empty() ? r.outdir + "/shots.bin" : r.shots_out;
    io::write_shots_raw(path, shots);
    std::printf("\nmodelling %.2f s; velocity %s\nwrote %s  (%d shots x %d traces x %d samples)\n", clock.elapsed(),
                r.velocity.empty() ? (r.outdir + "/velocity.bin").c_str() : r.velocity.c_str(), path.c_str(),
                r.nshots, shots.empty() ? 0 : shots.front().nrec(), r.nt);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception& e) {
        std::fprintf(stderr, "ERROR: %s\n", e.what());
        return 1;
    }
}
