// rtm_bandwidth_probe -- the bandwidth this host and its GPU deliver, the
// denominator of the roofline fractions.
//
// The command line and CSV of the reference's rtm_bandwidth_probe
// (src/tools/bandwidth_probe.cpp:90-124): --csv PATH --mb N --repeats N; one
// row per device appended to PATH with the columns
// date,host,device_kind,device_name,threads,gbps, which the reference's
// scripts/plot_benchmark.py:68-72 reads as the peak of each device.
//
//   cpu  a STREAM-style triad a[i] = b[i] + 3 c[i] over three buffers of
//        --mb MiB, split over every hardware thread (std::thread, one
//        contiguous range each, first touched by the thread that streams
//        it); 12 bytes per element moved; best of --repeats after one
//        warm-up pass.
//   gpu  rtm_hip_bandwidth_probe: a float4 copy between two device buffers
//        of --mb MiB, read + write bytes per pass, best of --repeats after
//        one warm-up, timed with HIP events.
// Exit status 0, or 1 with a message.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <sstream>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "cli.hpp"
#include "rtm/benchmark.hpp"
#include "rtm_hip.h"

namespace {

// Triad over [lo, hi) of each array.
void triad(float* a, const float* b, const float* c, size_t lo, size_t hi) {
    for (size_t i = lo; i < hi; ++i) a[i] = b[i] + 3.0f * c[i];
}

struct TriadResult {
    double gbps = 0.0;
    int threads = 1;
};

TriadResult cpu_triad(size_t elements, int repeats) {
    TriadResult r;
    r.threads = (int)std::max(1u, std::thread::hardware_concurrency());
    std::vector<float> a(elements), b(elements), c(elements);
    const size_t per = (elements + (size_t)r.threads - 1) / (size_t)r.threads;
    auto run = [&](auto&& body) {
        std::vector<std::thread> pool;
        for (int t = 0; t < r.threads; ++t) {
            const size_t lo = std::min(elements, (size_t)t * per), hi = std::min(elements, lo + per);
            pool.emplace_back([&body, lo, hi] { body(lo, hi); });
        }
        for (auto& th : pool) th.join();
    };
    // each thread initialises the range it streams
    run([&](size_t lo, size_t hi) {
        std::fill(b.begin() + (ptrdiff_t)lo, b.begin() + (ptrdiff_t)hi, 1.0f);
        std::fill(c.begin() + (ptrdiff_t)lo, c.begin() + (ptrdiff_t)hi, 2.0f);
        std::fill(a.begin() + (ptrdiff_t)lo, a.begin() + (ptrdiff_t)hi, 0.0f);
    });
    double best = 1e30;
    for (int pass = 0; pass <= repeats; ++pass) {
        const auto t0 = std::chrono::steady_clock::now();
        run([&](size_t lo, size_t hi) { triad(a.data(), b.data(), c.data(), lo, hi); });
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        if (pass > 0) best = std::min(best, s);
    }
    if (a[elements / 2] != 7.0f) throw std::runtime_error("triad produced a wrong value");
    r.gbps = 12.0 * (double)elements / best * 1e-9;
    return r;
}

std::string num(double v) {
    std::ostringstream o;
    o << v;
    return o.str();
}

std::string csv_safe(std::string s) {
    std::replace(s.begin(), s.end(), ',', ' ');
    return s;
}

void append_row(const std::string& path, const char* kind, const std::string& name, int threads, double gbps) {
    rtm::append_csv_row(path, "date,host,device_kind,device_name,threads,gbps",
                        {rtm::date_now_iso(), rtm::host_name(), kind, csv_safe(name), std::to_string(threads),
                         num(gbps)});
}

int run(int argc, char** argv) {
    std::string csv;
    int mb = 256, repeats = 5;
    rtm::cli::Options opts("usage: rtm_bandwidth_probe [--csv PATH] [--mb 256] [--repeats 5]");
    opts.section("Options")
        .text("--csv", csv, "append one row per device to PATH")
        .integer("--mb", mb, "MiB per buffer (default 256)")
        .integer("--repeats", repeats, "timed passes, best taken (default 5)");
    opts.action("--help", [&opts] { opts.usage(stdout); }, "this text");
    const auto rest = opts.parse(argc, argv);
    if (opts.stopped()) return 0;
    if (!rest.empty()) throw std::runtime_error("unknown option: " + rest.front());
    if (mb < 1 || repeats < 1) throw std::runtime_error("--mb and --repeats must be positive");
    const size_t bytes = (size_t)mb << 20;

    const TriadResult cpu = cpu_triad(bytes / sizeof(float), repeats);
    const std::string cpu_name = rtm::cpu_model_name();
    std::printf("CPU  %-40s  threads=%-3d  triad  %.1f GB/s\n", cpu_name.c_str(), cpu.threads, cpu.gbps);
    if (!csv.empty()) append_row(csv, "cpu", cpu_name, cpu.threads, cpu.gbps);

    if (rtm_hip_device_count() < 1) {
        std::printf("GPU  not probed: no HIP device\n");
        return 0;
    }
    std::string gpu_name = "unknown-gpu";
    rtm_hip_engine* h = nullptr;
    if (rtm_hip_create(0, &h) == RTM_HIP_OK) {
        rtm_hip_info info;
        if (rtm_hip_get_info(h, &info) == RTM_HIP_OK) gpu_name = info.device_name;
        rtm_hip_destroy(h);
    }
    double gbps = 0.0;
    if (rtm_hip_bandwidth_probe(0, bytes, repeats, &gbps) != RTM_HIP_OK) {
        std::printf("GPU  not probed: %s\n", rtm_hip_last_error(nullptr));
        return 0;
    }
    std::printf("GPU  %-40s  copy   %.1f GB/s\n", gpu_name.c_str(), gbps);
    if (!csv.empty()) append_row(csv, "gpu", gpu_name, 0, gbps);
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
