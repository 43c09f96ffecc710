// Three-phase benchmark driver (reference: include/lbm/benchmark/benchmark.hpp,
// src/lbm/benchmark/benchmark.cpp; run by `portlbm --benchmark`, the
// reference's BENCHMARK_MODE build, src/main_gpu.cpp:151-183).
//
// Phase 0 compiles nothing here (there is no JIT) but keeps its role as a
// short warm-up sweep of 10 steps; phase 1 runs every algorithm x data layout x
// work-group size of settings/benchmark.json on the Hagen-Poiseuille channel of
// the settings file and keeps, per algorithm, the layout and work-group size
// with the lowest mean initialisation + run time; phase 2 runs those on the
// phase-two scenarios.  Result files have the reference's names and schema
// (benchmarks/phase{0,1}/results_phase{p}_{algorithm}_{device}.json,
// benchmarks/phase2/phase2_{algorithm}_{scenario}_{device}.json) with four more
// arrays per configuration beside "initializationTimes" and "runtimes":
// "kernelTimes" (GPU seconds of the steps, HIP events), "mflups" (fluid-node
// updates per second / 1e6 on the GPU time), "gbps" (algorithmic bytes,
// 2 x 9 x sizeof(storage) per fluid-node update, per second / 1e9), and, once
// per file, "fluidNodes", "storage" and "bytesPerFluidNodeUpdate".
#ifndef LBM_BENCHMARK_HPP
#define LBM_BENCHMARK_HPP

#include <cstddef>
#include <filesystem>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../core/properties.hpp"
#include "../core/timer.hpp"
#include "../execution/hip_algorithm_handler.hpp"
#include "../file_interaction/json.hpp"

namespace lbm
{
namespace benchmark
{

// "Dd, Hh, Mm, Ss" (benchmark.hpp:40-54; the reference prints the day count in
// the hours field, this one prints the hours).
std::string seconds_to_time_format(unsigned int seconds);

// Progress over the test rows of one phase (benchmark.hpp:59-76, benchmark.cpp:13-26).
struct Progress
{
    int total_test_rows;
    int current_test_rows;
    double current_average_time;
    double progress;
    double remaining_time_estimation;

    explicit Progress(unsigned int total_test_rows);
    // one more row done, which took `time` seconds
    void advance(double time);
};

// Work-group sizes tried for a device maximum (benchmark.cpp:30-62): for each
// quarter i/4 of the maximum (i = 4..1), the sizes region - j (j = 0..3) that are
// powers of 2 or 4, plus the first even and the first odd other size.
std::vector<std::size_t> get_work_group_sizes(std::size_t max_work_group_size);

// The parsed settings/benchmark.json (benchmark.cpp:84-99).
struct BenchmarkSettings
{
    unsigned int test_runs = 10;
    bool test_linear_two_lattice = true;
    std::vector<std::string> data_layouts;
    std::vector<std::string> algorithms;
    std::vector<std::string> phase_2_scenarios;

    static BenchmarkSettings from_file(const std::string &path);
};

class Benchmark
{
  public:
    // The handler must have been created from a settings file: that file is
    // rewritten before every configuration and re-read by initialize(), as in
    // the reference (benchmark.cpp:142,302).  settings/benchmark.json is read
    // from the project root, the settings file's grandparent directory.
    explicit Benchmark(std::shared_ptr<execution::HipAlgorithmHandler> &handler);

    void phase_0();
    void phase_1();
    void phase_2();

    // Per algorithm, the (data layout, work-group size, mean total seconds)
    // phase 1 chose.
    const std::map<std::string, std::string> &optimal_data_layouts() const { return optimal_data_layouts_; }
    const std::map<std::string, unsigned int> &optimal_work_group_sizes() const { return optimal_work_group_sizes_; }
    const std::map<std::string, double> &optimal_total_times() const { return optimal_total_times_; }

  private:
    struct Run
    {
        double initialization_time, runtime, kernel_time, mflups, gbps;
    };

    Run run_once();
    file_interaction::JsonValue file_header(const std::string &algorithm) const;
    static void push_run(file_interaction::JsonValue &entry, const Run &r);
    void write(const std::string &path, const file_interaction::JsonValue &data) const;
    void phase_0_and_1_inner_loop(int phase, std::size_t work_group_size_index, const std::string &ofdir,
                                  file_interaction::JsonValue &file_data, Progress &progress);
    void phase_0_and_1_helper(int phase);
    void phase_2_inner_loop(const std::string &ofdir, file_interaction::JsonValue &file_data, Progress &progress);

    std::shared_ptr<execution::HipAlgorithmHandler> handler_;
    std::filesystem::path project_root_;
    std::string device_name_;
    std::unique_ptr<core::Properties> properties_;
    core::Timer runtime_timer_, benchmark_timer_;
    std::vector<std::size_t> work_group_sizes_;
    BenchmarkSettings settings_;
    std::map<std::string, std::string> optimal_data_layouts_;
    std::map<std::string, unsigned int> optimal_work_group_sizes_;
    std::map<std::string, double> optimal_total_times_;
};

}  // namespace benchmark
}  // namespace lbm

#endif  // LBM_BENCHMARK_HPP
