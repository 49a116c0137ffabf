// comparison.hpp -- image error metrics of the parity gate (same surface as
// the reference's include/comparison.hpp:9-25).  `ref` is the CPU image.
#pragma once
#include <cstddef>
#include <ostream>
#include <vector>

namespace rtm {

struct ErrorMetrics {
    std::size_t n = 0;
    double max_abs = 0;        // max |ref - test|
    double mean_abs = 0;       // mean |ref - test|
    double rms = 0;            // sqrt(mean (ref-test)^2)
    double l2_relative = 0;    // ||ref-test|| / ||ref||
    double max_relative = 0;   // max|ref-test| / max|ref|
    double correlation = 0;    // normalised zero-lag correlation
    double ref_max_abs = 0;
    double test_max_abs = 0;
    std::size_t argmax = 0;    // linear index of the worst sample
};

ErrorMetrics compare_images(const std::vector<float>& ref, const std::vector<float>& test);
void print_metrics(std::ostream& os, const ErrorMetrics& m, int nx = 0, int nz = 0);

}  // namespace rtm
