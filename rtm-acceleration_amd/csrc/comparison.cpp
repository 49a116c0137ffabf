// comparison.cpp -- the parity-gate metrics of two images (the surface of the
// reference's include/comparison.hpp:9-25, metrics as its
// src/benchmark/comparison.cpp:7-35 defines them).
//
// All sums run in double precision, sample by sample in index order, so the
// metrics are a property of the two images and not of this implementation:
//   max_abs, argmax   the largest |ref - test| and its first index
//   mean_abs, rms     mean of |ref - test| and of (ref - test)^2 (rooted)
//   l2_relative       ||ref - test|| / ||ref||           (0 for a zero ref)
//   max_relative      max_abs / max |ref|               (0 for a zero ref)
//   correlation       <ref, test> / (||ref|| ||test||)  (0 if either is 0)
#include "rtm/comparison.hpp"

#include <cmath>
#include <iomanip>
#include <stdexcept>

namespace rtm {

namespace {

// Running moments of the two images and of their difference.
struct Moments {
    double abs_diff = 0, sq_diff = 0;      // sum |d|, sum d^2
    double sq_ref = 0, sq_test = 0;        // sum r^2, sum t^2
    double dot = 0;                        // sum r t

    void add(double r, double t) {
        const double d = r - t;
        abs_diff += std::fabs(d);
        sq_diff += d * d;
        sq_ref += r * r;
        sq_test += t * t;
        dot += r * t;
    }
};

double ratio_or_zero(double num, double den) { return den > 0 ? num / den : 0.0; }

}  // namespace

ErrorMetrics compare_images(const std::vector<float>& ref, const std::vector<float>& test) {
    if (ref.size() != test.size()) throw std::runtime_error("image size mismatch in comparison");
    ErrorMetrics m;
    m.n = ref.size();
    if (m.n == 0) return m;

    Moments mo;
    for (std::size_t i = 0; i < m.n; ++i) {
        const double r = ref[i], t = test[i];
        mo.add(r, t);
        const double err = std::fabs(r - t);
        if (err > m.max_abs) {                 // strict: the first worst sample wins
            m.max_abs = err;
            m.argmax = i;
        }
        if (std::fabs(r) > m.ref_max_abs) m.ref_max_abs = std::fabs(r);
        if (std::fabs(t) > m.test_max_abs) m.test_max_abs = std::fabs(t);
    }
    const double n = static_cast<double>(m.n);
    m.mean_abs = mo.abs_diff / n;
    m.rms = std::sqrt(mo.sq_diff / n);
    m.l2_relative = mo.sq_ref > 0 ? std::sqrt(mo.sq_diff / mo.sq_ref) : 0.0;
    m.max_relative = ratio_or_zero(m.max_abs, m.ref_max_abs);
    m.correlation = ratio_or_zero(mo.dot, std::sqrt(mo.sq_ref * mo.sq_test));
    return m;
}

void print_metrics(std::ostream& os, const ErrorMetrics& m, int nx, int nz) {
    // one "label : value" row per metric, labels left-aligned in 24 columns
    auto row = [&os](const char* label) -> std::ostream& {
        return os << "  " << std::left << std::setw(24) << label << std::right << ": ";
    };
    const char* rule = "------------------------------------------\n";
    os << "\n---------- NUMERICAL COMPARISON ----------\n";
    row("samples") << m.n << "\n";
    const struct {
        const char* label;
        double value;
    } rows[] = {
        {"max |ref|", m.ref_max_abs},          {"max |test|", m.test_max_abs},
        {"max absolute error", m.max_abs},     {"mean absolute error", m.mean_abs},
        {"RMS error", m.rms},                  {"relative L2 error", m.l2_relative},
        {"max error / max |ref|", m.max_relative}, {"normalized correlation", m.correlation},
    };
    for (const auto& r : rows) row(r.label) << r.value << "\n";
    if (nx > 0 && nz > 0) {
        const std::size_t cols = static_cast<std::size_t>(nz);
        row("worst sample at (ix,iz)") << "(" << m.argmax / cols << ", " << m.argmax % cols << ")\n";
    }
    os << rule;
}

}  // namespace rtm
