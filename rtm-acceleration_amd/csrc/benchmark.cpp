// benchmark.cpp -- stage report and the benchmark CSV row (the reference's
// src/benchmark/benchmark.cpp:21-141).  Work is counted as the reference
// does: (nx+2nb)(nz+2nb) cells x nt steps x 2 propagations x nshots, at
// 6*half+5 flops per cell-step for the stencil.
#include "rtm/benchmark.hpp"

#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <stdexcept>

namespace rtm {

namespace {
std::string fmt_time(double s) {
    char b[64];
    if (s < 1e-3) std::snprintf(b, sizeof b, "%10.3f us", s * 1e6);
    else if (s < 1.0) std::snprintf(b, sizeof b, "%10.3f ms", s * 1e3);
    else std::snprintf(b, sizeof b, "%10.3f s ", s);
    return b;
}

double cell_steps(const BenchmarkContext& c) {
    return static_cast<double>(c.nx + 2 * c.nb) * static_cast<double>(c.nz + 2 * c.nb) * c.nt * 2.0 *
           c.nshots;
}

double stencil_gflop(const BenchmarkContext& c) {
    return cell_steps(c) * (6.0 * (c.order / 2) + 5.0) * 1e-9;
}

std::string no_commas(std::string s) {
    std::replace(s.begin(), s.end(), ',', ' ');
    return s;
}
}  // namespace

void print_benchmark_report(std::ostream& os, const BenchmarkContext& c) {
    const double gflop = stencil_gflop(c);
    os << "\n========================================\nRTM BENCHMARK\n========================================\n\n"
       << "Grid:\n"
       << "    NX             = " << c.nx << "  (dx = " << c.dx << " m)\n"
       << "    NZ             = " << c.nz << "  (dz = " << c.dz << " m)\n"
       << "    Sponge NB      = " << c.nb << "\n"
       << "    Extended       = " << (c.nx + 2 * c.nb) << " x " << (c.nz + 2 * c.nb) << "\n"
       << "    Spatial order  = " << c.order << "\n\n"
       << "Time steps:\n"
       << "    NT             = " << c.nt << "  (dt = " << c.dt * 1e3 << " ms, T = " << c.nt * c.dt << " s)\n\n"
       << "Shots:\n    NSHOTS         = " << c.nshots << "\n\n"
       << "Wavefield storage:\n    Snapshots      = " << (c.snapshot_bytes / (1024.0 * 1024.0)) << " MiB\n\n"
       << "Propagation work:\n    Stencil GFLOP  = " << gflop << "\n"
       << "    Cell-steps     = " << cell_steps(c) * 1e-9 << " G\n\n";
    const double base_total = c.results.empty() ? 0.0 : c.results[0].t.total;
    for (std::size_t k = 0; k < c.results.size(); ++k) {
        const auto& r = c.results[k];
        os << r.name << ":\n    Data loading:  " << fmt_time(r.t.io) << "\n";
        if (r.gpu)
            os << "    H2D:           " << fmt_time(r.t.h2d) << "\n"
               << "    Kernel:        " << fmt_time(r.t.kernel) << "\n"
               << "    D2H:           " << fmt_time(r.t.d2h) << "\n";
        os << "    Forward:       " << fmt_time(r.t.forward) << "\n"
           << "    Backward:      " << fmt_time(r.t.backward) << "   (imaging excluded)\n"
           << "    Imaging:       " << fmt_time(r.t.imaging) << (r.gpu ? "   (fused into backward)" : "") << "\n"
           << "    Total:         " << fmt_time(r.t.total) << "\n";
        const double prop = r.t.forward + r.t.backward;
        if (prop > 0.0)
            os << "    Stencil rate:  " << gflop / prop << " GFLOP/s\n"
               << "    Cell rate:     " << cell_steps(c) * 1e-9 / prop << " Gcell-steps/s\n";
        if (k > 0 && base_total > 0.0 && r.t.total > 0.0)
            os << "    Speedup:       " << base_total / r.t.total << " x\n";
        os << "\n";
    }
    os << "========================================\n";
}

std::string host_name() {
    char b[256] = {0};
    if (gethostname(b, sizeof b - 1) != 0) return "unknown-host";
    return b;
}

std::string cpu_model_name() {
    std::ifstream f("/proc/cpuinfo");
    std::string line;
    while (std::getline(f, line)) {
        if (line.rfind("model name", 0) == 0) {
            const auto colon = line.find(':');
            if (colon == std::string::npos) break;
            std::string name = line.substr(colon + 1);
            name.erase(0, name.find_first_not_of(' '));
            return name;
        }
    }
    return "unknown-cpu";
}

std::string date_now_iso() {
    const std::time_t now = std::time(nullptr);
    char b[32];
    std::strftime(b, sizeof b, "%Y-%m-%dT%H:%M:%S", std::localtime(&now));
    return b;
}

// Same 26 columns as the reference (src/benchmark/benchmark.cpp:107-111).
void append_benchmark_csv(const std::string& path, const BenchmarkContext& c, const EngineResult& r,
                          const RunInfo& info) {
    bool need_header = true;
    {
        std::ifstream existing(path);
        if (existing && existing.peek() != std::ifstream::traits_type::eof()) need_header = false;
    }
    std::ofstream out(path, std::ios::app);
    if (!out) throw std::runtime_error("cannot append benchmark csv: " + path);
    if (need_header)
        out << "date,host,cpu_name,threads,gpu_name,gpus,engine,dataset,nx,nz,nb,order,nt,dt,nshots,"
               "store_interval,t_io,t_h2d,t_forward,t_backward,t_imaging,t_d2h,t_total,stencil_gflop,"
               "stencil_gflops,snapshot_mib\n";
    const double gflop = stencil_gflop(c);
    const double prop = r.t.forward + r.t.backward;
    out << std::setprecision(9) << date_now_iso() << ',' << no_commas(info.host) << ','
        << no_commas(info.cpu_name) << ',' << info.threads << ',' << no_commas(info.gpu_name) << ','
        << info.gpus << ',' << no_commas(info.engine) << ',' << no_commas(info.dataset) << ',' << c.nx << ','
        << c.nz << ',' << c.nb << ',' << c.order << ',' << c.nt << ',' << c.dt << ',' << c.nshots << ','
        << c.store_interval << ',' << r.t.io << ',' << r.t.h2d << ',' << r.t.forward << ',' << r.t.backward
        << ',' << r.t.imaging << ',' << r.t.d2h << ',' << r.t.total << ',' << gflop << ','
        << (prop > 0 ? gflop / prop : 0.0) << ',' << (c.snapshot_bytes / (1024.0 * 1024.0)) << '\n';
}

}  // namespace rtm
