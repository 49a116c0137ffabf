// benchmark.cpp -- the stage report and the benchmark CSV row (the surface of
// the reference's include/benchmark.hpp:12-67).
//
// Work is counted as the reference counts it (src/benchmark/benchmark.cpp:
// 21-25): every extended cell (nx+2nb)(nz+2nb) once per time step, for the
// forward and the backward propagation of every shot, at 6*half+5 stencil
// flops per cell-step.  The CSV keeps the reference's 26 columns and their
// number format, which scripts/pod/summary.py and scripts/plot_benchmark.py
// read.
#include "rtm/benchmark.hpp"

#include <sys/utsname.h>

#include <cstdio>
#include <ctime>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace rtm {

namespace {

struct Work {
    double cell_steps;   // extended cells x steps x 2 propagations x shots
    double gflop;        // stencil flops of those cell-steps, x 1e-9
};

Work work_of(const BenchmarkContext& c) {
    Work w;
    w.cell_steps = double(c.nx + 2 * c.nb) * double(c.nz + 2 * c.nb) * double(c.nt) * 2.0 * double(c.nshots);
    w.gflop = w.cell_steps * (6.0 * (c.order / 2) + 5.0) * 1e-9;
    return w;
}

// A duration with the unit that keeps three to four significant digits.
std::string duration(double s) {
    const struct {
        double scale;
        const char* unit;
    } units[] = {{1e6, "us"}, {1e3, "ms"}, {1.0, "s "}};
    const auto& u = s < 1e-3 ? units[0] : (s < 1.0 ? units[1] : units[2]);
    char b[48];
    std::snprintf(b, sizeof b, "%10.3f %s", s * u.scale, u.unit);
    return b;
}

// CSV fields never contain the separator.
std::string field(std::string s) {
    for (char& ch : s)
        if (ch == ',') ch = ' ';
    return s;
}

// A number as the CSV writes it: an ostream at 9 significant digits.
template <typename T>
std::string num(T v) {
    std::ostringstream o;
    o << std::setprecision(9) << v;
    return o.str();
}

const char* const kCsvColumns[] = {
    "date",   "host",           "cpu_name", "threads",   "gpu_name",  "gpus",      "engine",
    "dataset", "nx",            "nz",       "nb",        "order",     "nt",        "dt",
    "nshots", "store_interval", "t_io",     "t_h2d",     "t_forward", "t_backward", "t_imaging",
    "t_d2h",  "t_total",        "stencil_gflop", "stencil_gflops", "snapshot_mib"};

std::string join(const std::vector<std::string>& v) {
    std::string s;
    for (std::size_t i = 0; i < v.size(); ++i) s += (i ? "," : "") + v[i];
    return s;
}

}  // namespace

void append_csv_row(const std::string& path, const std::string& header, const std::vector<std::string>& fields) {
    bool empty = true;
    {
        std::ifstream in(path, std::ios::binary);
        empty = !in || in.peek() == std::ifstream::traits_type::eof();
    }
    std::ofstream out(path, std::ios::app);
    if (!out) throw std::runtime_error("cannot append csv: " + path);
    if (empty) out << header << '\n';
    out << join(fields) << '\n';
}

void print_benchmark_report(std::ostream& os, const BenchmarkContext& c) {
    const Work w = work_of(c);
    const std::string bar(40, '=');
    auto kv = [&os](const char* key) -> std::ostream& { return os << "    " << std::left << std::setw(15) << key
                                                                   << std::right << "= "; };
    os << "\n" << bar << "\nRTM BENCHMARK\n" << bar << "\n\nGrid:\n";
    kv("NX") << c.nx << "  (dx = " << c.dx << " m)\n";
    kv("NZ") << c.nz << "  (dz = " << c.dz << " m)\n";
    kv("Sponge NB") << c.nb << "\n";
    kv("Extended") << (c.nx + 2 * c.nb) << " x " << (c.nz + 2 * c.nb) << "\n";
    kv("Spatial order") << c.order << "\n\nTime steps:\n";
    kv("NT") << c.nt << "  (dt = " << c.dt * 1e3 << " ms, T = " << c.nt * c.dt << " s)\n\nShots:\n";
    kv("NSHOTS") << c.nshots << "\n\nWavefield storage:\n";
    kv("Snapshots") << c.snapshot_bytes / (1024.0 * 1024.0) << " MiB\n\nPropagation work:\n";
    kv("Stencil GFLOP") << w.gflop << "\n";
    kv("Cell-steps") << w.cell_steps * 1e-9 << " G\n\n";

    auto line = [&os](const char* stage, double s, const char* note = "") {
        os << "    " << std::left << std::setw(15) << (std::string(stage) + ":") << std::right << duration(s) << note
           << "\n";
    };
    const double baseline = c.results.empty() ? 0.0 : c.results.front().t.total;
    for (std::size_t k = 0; k < c.results.size(); ++k) {
        const EngineResult& r = c.results[k];
        os << r.name << ":\n";
        line("Data loading", r.t.io);
        if (r.gpu) {
            line("H2D", r.t.h2d);
            line("Kernel", r.t.kernel);
            line("D2H", r.t.d2h);
        }
        line("Forward", r.t.forward);
        line("Backward", r.t.backward, "   (imaging excluded)");
        line("Imaging", r.t.imaging, r.gpu ? "   (fused into backward)" : "");
        line("Total", r.t.total);
        const double prop = r.t.forward + r.t.backward;
        if (prop > 0.0) {
            os << "    Stencil rate:  " << w.gflop / prop << " GFLOP/s\n";
            os << "    Cell rate:     " << w.cell_steps * 1e-9 / prop << " Gcell-steps/s\n";
        }
        if (k > 0 && baseline > 0.0 && r.t.total > 0.0) os << "    Speedup:       " << baseline / r.t.total << " x\n";
        os << "\n";
    }
    os << bar << "\n";
}

std::string host_name() {
    struct utsname u;
    if (uname(&u) != 0 || u.nodename[0] == '\0') return "unknown-host";
    return u.nodename;
}

std::string cpu_model_name() {
    std::ifstream in("/proc/cpuinfo");
    for (std::string line; std::getline(in, line);) {
        const auto colon = line.find(':');
        if (colon == std::string::npos) continue;
        std::string key = line.substr(0, colon);
        key.erase(key.find_last_not_of(" \t") + 1);
        if (key != "model name") continue;
        const auto start = line.find_first_not_of(' ', colon + 1);
        return start == std::string::npos ? std::string() : line.substr(start);
    }
    return "unknown-cpu";
}

std::string date_now_iso() {
    const std::time_t now = std::time(nullptr);
    std::tm local{};
    localtime_r(&now, &local);
    std::ostringstream o;
    o << std::put_time(&local, "%Y-%m-%dT%H:%M:%S");
    return o.str();
}

void append_benchmark_csv(const std::string& path, const BenchmarkContext& c, const EngineResult& r,
                          const RunInfo& info) {
    const Work w = work_of(c);
    const double prop = r.t.forward + r.t.backward;
    std::vector<std::string> header(std::begin(kCsvColumns), std::end(kCsvColumns));
    const std::vector<std::string> row = {
        date_now_iso(), field(info.host), field(info.cpu_name), num(info.threads), field(info.gpu_name),
        num(info.gpus), field(info.engine), field(info.dataset), num(c.nx), num(c.nz), num(c.nb), num(c.order),
        num(c.nt), num(c.dt), num(c.nshots), num(c.store_interval), num(r.t.io), num(r.t.h2d), num(r.t.forward),
        num(r.t.backward), num(r.t.imaging), num(r.t.d2h), num(r.t.total), num(w.gflop),
        num(prop > 0 ? w.gflop / prop : 0.0), num(c.snapshot_bytes / (1024.0 * 1024.0))};
    append_csv_row(path, join(header), row);
}

}  // namespace rtm
