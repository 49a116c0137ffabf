// rtm_compare -- the parity gate between two float32 images.
//
// usage: rtm_compare REFERENCE TEST [--nx N --nz N] [--csv PATH --engine NAME
// --dataset NAME]: prints the metrics of TEST against REFERENCE (the CPU
// image), passes when the relative L2 error is below 1e-5 and the normalised
// correlation above 0.9999 (src/tools/compare_main.cpp:52), and optionally
// appends one row to a CSV with the reference's 11 columns, which
// scripts/pod/summary.py reads.  Exit status 0 pass, 2 fail, 1 error.
#include <cstdio>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "cli.hpp"
#include "rtm/benchmark.hpp"
#include "rtm/comparison.hpp"
#include "rtm/rtm_io.hpp"

using namespace rtm;

namespace {

struct Gate {
    double max_l2 = 1e-5;          // relative L2 error below this
    double min_correlation = 0.9999;
    bool passes(const ErrorMetrics& m) const { return m.l2_relative < max_l2 && m.correlation > min_correlation; }
};

std::string fmt9(double v) {
    std::ostringstream o;
    o.precision(9);
    o << v;
    return o.str();
}

int run(int argc, char** argv) {
    int nx = 0, nz = 0;
    std::string csv, engine = "unknown", dataset = "unknown";
    cli::Options opts("usage: rtm_compare REFERENCE.bin TEST.bin [--nx N --nz N]\n"
                      "                   [--csv PATH --engine NAME --dataset NAME]\n"
                      "  REFERENCE is the CPU reference image; exit 0 pass, 2 fail, 1 error.");
    opts.section("Options")
        .integer("--nx", nx, "image samples along x (to locate the worst sample)")
        .integer("--nz", nz, "image samples along z")
        .text("--csv", csv, "append one row (metrics and pass/fail) to PATH")
        .text("--engine", engine, "engine label of the CSV row")
        .text("--dataset", dataset, "dataset label of the CSV row");
    const std::vector<std::string> files = opts.parse(argc, argv);
    if (files.size() != 2) {
        opts.usage(stdout);
        if (files.size() > 2) throw std::runtime_error("unknown option: " + files[2]);
        return 1;
    }
    const ErrorMetrics m = compare_images(io::read_raw_floats(files[0]), io::read_raw_floats(files[1]));
    print_metrics(std::cout, m, nx, nz);

    const Gate gate;
    const bool pass = gate.passes(m);
    std::printf("\nDefault gate: L2rel < %.0e and corr > %.4f  =>  %s\n", gate.max_l2, gate.min_correlation,
                pass ? "PASS" : "FAIL");
    if (!csv.empty()) {
        append_csv_row(csv, "date,host,engine,dataset,max_abs,mean_abs,rms,l2_relative,max_relative,correlation,pass",
                       {date_now_iso(), host_name(), engine, dataset, fmt9(m.max_abs), fmt9(m.mean_abs), fmt9(m.rms),
                        fmt9(m.l2_relative), fmt9(m.max_relative), fmt9(m.correlation), pass ? "1" : "0"});
        std::printf("appended compare row to %s\n", csv.c_str());
    }
    return pass ? 0 : 2;
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
