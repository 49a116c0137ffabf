// benchmark.hpp -- wall clocks and the stage report (same surface as the
// reference's include/benchmark.hpp:12-67).
#pragma once
#include <chrono>
#include <ostream>
#include <string>
#include <vector>
#include "rtm_engine.hpp"

namespace rtm {

class Timer {
public:
    Timer() { reset(); }
    void reset() { t0_ = clock::now(); }
    double elapsed() const { return std::chrono::duration<double>(clock::now() - t0_).count(); }
private:
    using clock = std::chrono::steady_clock;
    clock::time_point t0_;
};

class ScopedAccumulator {
public:
    explicit ScopedAccumulator(double& acc) : acc_(acc) {}
    ~ScopedAccumulator() { acc_ += t_.elapsed(); }
private:
    double& acc_;
    Timer t_;
};

struct EngineResult {
    std::string name;
    StageTimes t;
    bool gpu = false;
};

struct BenchmarkContext {
    int nx = 0, nz = 0, nb = 0, nt = 0, nshots = 0, order = 0;
    int store_interval = 1;
    float dx = 0, dz = 0, dt = 0;
    std::size_t snapshot_bytes = 0;
    std::vector<EngineResult> results;   // results[0] is the baseline
};

void print_benchmark_report(std::ostream& os, const BenchmarkContext& ctx);

struct RunInfo {
    std::string host, cpu_name, gpu_name, engine, dataset;
    int threads = 1;
    int gpus = 0;
};

std::string host_name();
std::string cpu_model_name();
std::string date_now_iso();

void append_benchmark_csv(const std::string& path, const BenchmarkContext& ctx,
                          const EngineResult& result, const RunInfo& info);

// Appends one row of comma-joined fields to the CSV at `path`, writing
// `header` first when the file is new or empty (the benchmark and the
// comparison CSVs both use it).
void append_csv_row(const std::string& path, const std::string& header, const std::vector<std::string>& fields);

}  // namespace rtm
