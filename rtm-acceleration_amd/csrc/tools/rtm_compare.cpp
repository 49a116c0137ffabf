// rtm_compare -- the parity gate (the reference's src/tools/compare_main.cpp:13-82):
// metrics of TEST against REFERENCE, gate L2rel < 1e-5 and correlation > 0.9999,
// exit 0 pass / 2 fail / 1 error, optional CSV row.
#include <cstdio>
#include <fstream>
#include <iostream>
#include <stdexcept>
#include <string>

#include "rtm/benchmark.hpp"
#include "rtm/comparison.hpp"
#include "rtm/rtm_io.hpp"

using namespace rtm;

int main(int argc, char** argv) try {
    if (argc < 3) {
        std::printf("usage: rtm_compare REFERENCE.bin TEST.bin [--nx N --nz N]\n"
                    "                   [--csv PATH --engine NAME --dataset NAME]\n");
        return 1;
    }
    const std::string ref_path = argv[1], test_path = argv[2];
    int nx = 0, nz = 0;
    std::string csv_path, engine = "unknown", dataset = "unknown";
    for (int i = 3; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "--nx" && i + 1 < argc) nx = std::stoi(argv[++i]);
        else if (a == "--nz" && i + 1 < argc) nz = std::stoi(argv[++i]);
        else if (a == "--csv" && i + 1 < argc) csv_path = argv[++i];
        else if (a == "--engine" && i + 1 < argc) engine = argv[++i];
        else if (a == "--dataset" && i + 1 < argc) dataset = argv[++i];
        else throw std::runtime_error("unknown option: " + a);
    }
    const auto m = compare_images(io::read_raw_floats(ref_path), io::read_raw_floats(test_path));
    print_metrics(std::cout, m, nx, nz);
    const double gate_l2 = 1e-5, gate_corr = 0.9999;
    const bool pass = m.l2_relative < gate_l2 && m.correlation > gate_corr;
    std::printf("\nDefault gate: L2rel < %.0e and corr > %.4f  =>  %s\n", gate_l2, gate_corr,
                pass ? "PASS" : "FAIL");
    if (!csv_path.empty()) {
        bool header = true;
        {
            std::ifstream existing(csv_path);
            if (existing && existing.peek() != std::ifstream::traits_type::eof()) header = false;
        }
        std::ofstream out(csv_path, std::ios::app);
        if (!out) throw std::runtime_error("cannot append compare csv: " + csv_path);
        if (header) out << "date,host,engine,dataset,max_abs,mean_abs,rms,l2_relative,max_relative,correlation,pass\n";
        out.precision(9);
        out << date_now_iso() << ',' << host_name() << ',' << engine << ',' << dataset << ',' << m.max_abs << ','
            << m.mean_abs << ',' << m.rms << ',' << m.l2_relative << ',' << m.max_relative << ','
            << m.correlation << ',' << (pass ? 1 : 0) << '\n';
    }
    return pass ? 0 : 2;
} catch (const std::exception& e) {
    std::fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
}
