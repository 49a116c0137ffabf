// rtm_synth -- synthetic shot gathers by forward modelling on the GPU.
//
// Same flags, built-in model, acquisition geometry and outputs as the
// reference's `rtm_synth` (src/tools/make_synthetic.cpp:28-177): the model is
// either --velocity PATH (with its .hdr) or the built-in three layers with a
// dipping interface and a box diffractor; receivers every --rec-dx along the
// line; shots spread evenly or at --shot-x0 + i*--shot-dx; every shot
// modelled with model_shot and written as RTMS.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtm/benchmark.hpp"
#include "rtm/rtm_factory.hpp"
#include "rtm/rtm_io.hpp"

using namespace rtm;

int main(int argc, char** argv) try {
    std::string outdir = "data/synthetic", vel_path, shots_out, engine_name = "hip";
    int nx = 201, nz = 151, nt = 1400, nshots = 4, order = 8, nb = 60;
    float dx = 10.0f, dz = 10.0f, dt = 0.001f, f0 = 12.0f;
    float src_z = 20.0f, rec_z = 20.0f, rec_dx = 20.0f, shot_x0 = -1.0f, shot_dx = -1.0f;
    io::VelocityHeader vh;
    bool set_nx = false, set_nz = false, set_dx = false, set_dz = false;

    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) throw std::runtime_error("missing value after " + a);
            return argv[++i];
        };
        if (a == "--outdir") outdir = val();
        else if (a == "--nx") { nx = std::stoi(val()); set_nx = true; }
        else if (a == "--nz") { nz = std::stoi(val()); set_nz = true; }
        else if (a == "--dx") { dx = std::stof(val()); set_dx = true; }
        else if (a == "--dz") { dz = std::stof(val()); set_dz = true; }
        else if (a == "--nt") nt = std::stoi(val());
        else if (a == "--dt") dt = std::stof(val());
        else if (a == "--f0") f0 = std::stof(val());
        else if (a == "--nshots") nshots = std::stoi(val());
        else if (a == "--order") order = std::stoi(val());
        else if (a == "--nb") nb = std::stoi(val());
        else if (a == "--rec-dx") rec_dx = std::stof(val());
        else if (a == "--src-z") src_z = std::stof(val());
        else if (a == "--rec-z") rec_z = std::stof(val());
        else if (a == "--velocity") vel_path = val();
        else if (a == "--vel-layout") vh.layout = val();
        else if (a == "--shots-out") shots_out = val();
        else if (a == "--shot-x0") shot_x0 = std::stof(val());
        else if (a == "--shot-dx") shot_dx = std::stof(val());
        else if (a == "--engine") engine_name = val();
        else throw std::runtime_error("unknown option: " + a);
    }

    VelocityModel model;
    if (!vel_path.empty()) {
        io::VelocityHeader side;
        if (io::read_header_file(vel_path + ".hdr", side)) {
            if (!set_nx) nx = side.nx;
            if (!set_nz) nz = side.nz;
            if (!set_dx) dx = side.dx;
            if (!set_dz) dz = side.dz;
            if (vh.layout == "zfast") vh.layout = side.layout;
            vh.ox = side.ox;
            vh.oz = side.oz;
        }
        vh.nx = nx; vh.nz = nz; vh.dx = dx; vh.dz = dz;
        model = io::read_velocity_raw(vel_path, vh);
    } else {
        // three layers: 1500 m/s, 2000 m/s below a flat interface at 400 m,
        // 2500 m/s below a dipping interface, and a 60 x 60 m box at 2800 m/s
        model.grid.nx = nx; model.grid.nz = nz; model.grid.dx = dx; model.grid.dz = dz;
        model.v.assign(static_cast<std::size_t>(nx) * nz, 1500.0f);
        for (int ix = 0; ix < nx; ++ix) {
            const float x = ix * dx;
            const float z_dip = 900.0f + 0.15f * (x - 0.5f * nx * dx);
            for (int iz = 0; iz < nz; ++iz) {
                const float z = iz * dz;
                float v = 1500.0f;
                if (z >= 400.0f) v = 2000.0f;
                if (z >= z_dip) v = 2500.0f;
                if (x >= 1200.0f && x < 1260.0f && z >= 700.0f && z < 760.0f) v = 2800.0f;
                model.v[static_cast<std::size_t>(ix) * nz + iz] = v;
            }
        }
        vh.nx = nx; vh.nz = nz; vh.dx = dx; vh.dz = dz; vh.ox = 0; vh.oz = 0; vh.layout = "zfast";
        io::write_velocity_raw(outdir + "/velocity.bin", model);
        io::write_header_file(outdir + "/velocity.bin.hdr", vh);
    }

    // split-spread surface line, every receiver live for every shot
    const int nrec = static_cast<int>((nx - 1) * dx / rec_dx) + 1;
    const float xmax = (nx - 1) * dx;
    std::vector<ShotRecord> shots(static_cast<std::size_t>(nshots));
    for (int is = 0; is < nshots; ++is) {
        ShotRecord& s = shots[is];
        s.sx = shot_dx > 0.0f ? (shot_x0 >= 0.0f ? shot_x0 : 0.0f) + shot_dx * static_cast<float>(is)
                              : xmax * static_cast<float>(is + 1) / static_cast<float>(nshots + 1);
        s.sz = src_z;
        s.nt = nt;
        s.dt = dt;
        s.rx.resize(nrec);
        s.rz.assign(nrec, rec_z);
        for (int ir = 0; ir < nrec; ++ir) s.rx[ir] = ir * rec_dx;
    }

    RTMParams par;
    par.nb = nb; par.order = order; par.f0 = f0; par.store_interval = 1;
    TimeAxis ta;
    ta.nt = nt; ta.dt = dt;
    auto engine = make_engine(engine_name);
    engine->setup(model, par, ta);
    const float dt_max = max_stable_dt(model.vmax(), engine->grid(), order);
    const float ppw = points_per_wavelength(model.vmin(), f0, engine->grid());
    std::printf("MODELLED SHOT GATHERS (inverse crime: same propagator as the migration)\n"
                "  grid %d x %d, dx = %.1f m, dz = %.1f m\n  vmin = %.0f, vmax = %.0f m/s\n"
                "  dt = %.6f s (max stable %.6f s), nt = %d, T = %.2f s\n"
                "  points per wavelength = %.2f\n  %d shots, %d receivers each\n\n",
                nx, nz, dx, dz, model.vmin(), model.vmax(), dt, dt_max, nt, nt * dt, ppw, nshots, nrec);
    if (dt > dt_max) throw std::runtime_error("chosen dt is unstable");

    Timer t;
    for (int is = 0; is < nshots; ++is) {
        std::printf("  modelling shot %d / %d (sx = %.1f m) ...\n", is + 1, nshots, shots[is].sx);
        std::fflush(stdout);
        engine->model_shot(shots[is]);
    }
    const std::string path = shots_out.empty() ? outdir + "/shots.bin" : shots_out;
    io::write_shots_raw(path, shots);
    std::printf("\nmodelling took %.2f s\nvelocity model: %s\nwrote %s  (%d shots x %d traces x %d samples)\n",
                t.elapsed(), vel_path.empty() ? (outdir + "/velocity.bin").c_str() : vel_path.c_str(),
                path.c_str(), nshots, nrec, nt);
    return 0;
} catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
}
