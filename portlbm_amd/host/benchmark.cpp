// Three-phase benchmark driver (reference: src/lbm/benchmark/benchmark.cpp).
#include "lbm/benchmark/benchmark.hpp"

#include <cmath>
#include <fstream>
#include <iostream>
#include <limits>
#include <numeric>
#include <sstream>

#include "lbm/core/domain.hpp"
#include "lbm/exceptions/exceptions.hpp"
#include "lbm/file_interaction/file_interaction.hpp"

namespace lbm
{
namespace benchmark
{

using file_interaction::JsonValue;

std::string seconds_to_time_format(unsigned int seconds)
{
    const unsigned int days = seconds / 86400;
    seconds -= days * 86400;
    const unsigned int hours = seconds / 3600;
    seconds -= hours * 3600;
    const unsigned int minutes = seconds / 60;
    seconds -= minutes * 60;
    return std::to_string(days) + "d, " + std::to_string(hours) + "h, " + std::to_string(minutes) + "m, " +
           std::to_string(seconds) + "s";
}

Progress::Progress(unsigned int total) :
    total_test_rows(static_cast<int>(total)),
    current_test_rows(0),
    current_average_time(0.0),
    progress(0.0),
    remaining_time_estimation(0.0)
{
}

void Progress::advance(double time)
{
    ++current_test_rows;
    progress = 100.0 * static_cast<double>(current_test_rows) / static_cast<double>(total_test_rows);
    current_average_time = ((current_test_rows - 1) * current_average_time + time) / current_test_rows;
    remaining_time_estimation = static_cast<double>(total_test_rows - current_test_rows) * current_average_time;
}

std::vector<std::size_t> get_work_group_sizes(std::size_t max_work_group_size)
{
    std::vector<std::size_t> sizes;
    for (std::size_t i = 4; i >= 1; --i)
    {
        const std::size_t region = i * max_work_group_size / 4;
        bool even_taken = false, odd_taken = false;
        for (std::size_t j = 0; j < 4; ++j)
        {
            const std::size_t size = region - j;
            if (core::power_functions::which_power_of_4(size) || core::power_functions::which_power_of_2(size))
                sizes.push_back(size);
            else if (size % 2 == 0 && !even_taken)
            {
                sizes.push_back(size);
                even_taken = true;
            }
            else if (size % 2 == 1 && !odd_taken)
            {
                sizes.push_back(size);
                odd_taken = true;
            }
        }
    }
    return sizes;
}

namespace
{
std::vector<std::string> string_list(const JsonValue &v, const std::string &what)
{
    std::vector<std::string> out;
    for (const JsonValue &e : v.items())
        out.push_back(e.as_string(what));
    return out;
}

unsigned int storage_bytes(const std::string &storage)
{
    if (storage == "f64")
        return 8;
    if (storage == "f32")
        return 4;
    return 2;
}
}  // namespace

BenchmarkSettings BenchmarkSettings::from_file(const std::string &path)
{
    std::ifstream in(path);
    if (!in)
        throw exceptions::Exception("cannot read benchmark settings \"" + path + "\"");
    std::stringstream ss;
    ss << in.rdbuf();
    const JsonValue doc = JsonValue::parse(ss.str());
    BenchmarkSettings s;
    s.test_runs = static_cast<unsigned int>(doc.at("testRunsPerConfig").as_unsigned("testRunsPerConfig"));
    s.test_linear_two_lattice = doc.at("testLinearTwoLattice").as_bool("testLinearTwoLattice");
    s.data_layouts = string_list(doc.at("dataLayouts"), "dataLayouts");
    s.algorithms = string_list(doc.at("algorithms"), "algorithms");
    s.phase_2_scenarios = string_list(doc.at("phaseTwoScenarios"), "phaseTwoScenarios");
    if (s.test_runs == 0)
        throw exceptions::json::PropertyArgumentException("testRunsPerConfig must be positive");
    return s;
}

Benchmark::Benchmark(std::shared_ptr<execution::HipAlgorithmHandler> &handler) :
    handler_(handler),
    project_root_(std::filesystem::path(handler->get_settings_path()).parent_path().parent_path()),
    device_name_(execution::HipAlgorithmHandler::device_name()),
    properties_(std::make_unique<core::Properties>(
        file_interaction::json_to_properties(handler->get_settings_path(), -2))),
    work_group_sizes_(get_work_group_sizes(handler->processing_element_constraint)),
    settings_(BenchmarkSettings::from_file((project_root_ / "settings" / "benchmark.json").string()))
{
    properties_->debug_mode = false;
    properties_->frame_update_interval = properties_->time_steps + 1;
    properties_->scenario = "Hagen-Poiseuille";
    for (const std::string &a : settings_.algorithms)
    {
        optimal_data_layouts_[a] = "stream";
        optimal_work_group_sizes_[a] = 1;
        optimal_total_times_[a] = std::numeric_limits<double>::max();
    }
    if (settings_.test_linear_two_lattice)
    {
        optimal_data_layouts_["lptl"] = "stream";
        optimal_work_group_sizes_["lptl"] = 1;
        optimal_total_times_["lptl"] = std::numeric_limits<double>::max();
    }
}

Benchmark::Run Benchmark::run_once()
{
    handler_->initialize();
    Run r{};
    r.initialization_time = handler_->initialization_time;
    runtime_timer_.restart();
    handler_->start();
    handler_->block_until_finished();
    r.runtime = runtime_timer_.elapsed();
    r.kernel_time = handler_->device_run_seconds();
    const double updates = static_cast<double>(handler_->fluid_node_count()) * handler_->get_current_iteration();
    r.mflups = r.kernel_time > 0 ? updates / r.kernel_time / 1e6 : 0.0;
    r.gbps = r.mflups * 1e6 * 18.0 * storage_bytes(properties_->extensions.storage) / 1e9;
    return r;
}

JsonValue Benchmark::file_header(const std::string &algorithm) const
{
    JsonValue d = JsonValue::object();
    d["device"] = JsonValue::string(device_name_);
    d["algorithm"] = JsonValue::string(algorithm);
    d["verticalNodes"] = JsonValue::unsigned_number(properties_->vertical_nodes);
    d["horizontalNodes"] = JsonValue::unsigned_number(properties_->horizontal_nodes);
    d["storage"] = JsonValue::string(properties_->extensions.storage);
    d["bytesPerFluidNodeUpdate"] = JsonValue::unsigned_number(18u * storage_bytes(properties_->extensions.storage));
    return d;
}

void Benchmark::push_run(JsonValue &entry, const Run &r)
{
    entry["initializationTimes"].push_back(JsonValue::number(r.initialization_time));
    entry["runtimes"].push_back(JsonValue::number(r.runtime));
    entry["kernelTimes"].push_back(JsonValue::number(r.kernel_time));
    entry["mflups"].push_back(JsonValue::number(r.mflups));
    entry["gbps"].push_back(JsonValue::number(r.gbps));
}

void Benchmark::write(const std::string &path, const JsonValue &data) const
{
    std::ofstream file(path);
    if (!file)
        throw exceptions::Exception("cannot write benchmark results \"" + path + "\"");
    file << data.dump(4) << "\n";
}

namespace
{
JsonValue run_arrays(JsonValue entry)
{
    for (const char *k : {"initializationTimes", "runtimes", "kernelTimes", "mflups", "gbps"})
        entry[k] = JsonValue::array();
    return entry;
}

void print_progress(const Progress &p)
{
    std::cout << "Progress: " << p.current_test_rows << " / " << p.total_test_rows << " ("
              << std::floor(p.progress) << "%), remaining time estimation: "
              << seconds_to_time_format(static_cast<unsigned int>(p.remaining_time_estimation)) << " \n"
              << std::flush;
}
}  // namespace

// benchmark.cpp:171-231
void Benchmark::phase_0_and_1_inner_loop(int phase, std::size_t index, const std::string &ofdir,
                                         JsonValue &file_data, Progress &progress)
{
    for (const std::string &layout : settings_.data_layouts)
    {
        std::vector<double> total_times;
        properties_->data_layout = layout;
        file_interaction::properties_to_json(*properties_);
        for (unsigned int t = 0; t < settings_.test_runs; ++t)
        {
            benchmark_timer_.restart();
            const Run r = run_once();
            push_run(file_data[layout].item(index), r);
            progress.advance(benchmark_timer_.elapsed());
            print_progress(progress);
            write(ofdir, file_data);
            total_times.push_back(r.initialization_time + r.runtime);
        }
        const double mean = std::accumulate(total_times.begin(), total_times.end(), 0.0) / settings_.test_runs;
        if (phase == 1 && mean < optimal_total_times_[properties_->algorithm])
        {
            optimal_data_layouts_[properties_->algorithm] = properties_->data_layout;
            optimal_work_group_sizes_[properties_->algorithm] = properties_->work_group_size;
            optimal_total_times_[properties_->algorithm] = mean;
        }
    }
}

// benchmark.cpp:233-283
void Benchmark::phase_0_and_1_helper(int phase)
{
    std::cout << std::string(100, '=') << "\nStarting benchmark phase " << phase << ".\n\n";
    unsigned int total = static_cast<unsigned int>(settings_.algorithms.size() * settings_.data_layouts.size() *
                                                   work_group_sizes_.size() * settings_.test_runs);
    if (settings_.test_linear_two_lattice)
        total += static_cast<unsigned int>(settings_.data_layouts.size() * settings_.test_runs);
    Progress progress(total);
    const std::string dir = (project_root_ / "benchmarks" / ("phase" + std::to_string(phase))).string();
    for (const std::string &algorithm : settings_.algorithms)
    {
        properties_->algorithm = algorithm;
        const std::string ofdir = dir + "/results_phase" + std::to_string(phase) + "_" + algorithm + "_" +
                                  device_name_ + ".json";
        JsonValue file_data = file_header(algorithm);
        for (const std::string &layout : settings_.data_layouts)
        {
            file_data[layout] = JsonValue::array();
            for (std::size_t wg : work_group_sizes_)
            {
                JsonValue e = JsonValue::object();
                e["workGroupSize"] = JsonValue::unsigned_number(wg);
                file_data[layout].push_back(run_arrays(e));
            }
        }
        for (std::size_t j = 0; j < work_group_sizes_.size(); ++j)
        {
            properties_->work_group_size = static_cast<unsigned int>(work_group_sizes_[j]);
            phase_0_and_1_inner_loop(phase, j, ofdir, file_data, progress);
        }
    }
    if (settings_.test_linear_two_lattice)
    {
        // lptl launches a linear range; "auto" in the reference, 256 here
        properties_->algorithm = "lptl";
        properties_->work_group_size = 256;
        const std::string ofdir =
            dir + "/results_phase" + std::to_string(phase) + "_lptl_" + device_name_ + ".json";
        JsonValue file_data = file_header("lptl");
        for (const std::string &layout : settings_.data_layouts)
        {
            file_data[layout] = JsonValue::array();
            JsonValue e = JsonValue::object();
            e["workGroupSize"] = JsonValue::string("auto");
            file_data[layout].push_back(run_arrays(e));
        }
        phase_0_and_1_inner_loop(phase, 0, ofdir, file_data, progress);
    }
    std::cout << "\nSuccessfully completed phase " << phase << " of benchmark\n" << std::string(100, '=') << "\n\n";
}

// benchmark.cpp:285-291: a short sweep (10 steps) before the measured one
void Benchmark::phase_0()
{
    const unsigned int steps = properties_->time_steps;
    properties_->time_steps = 10;
    properties_->frame_update_interval = 11;
    phase_0_and_1_helper(0);
    properties_->time_steps = steps;
    properties_->frame_update_interval = steps + 1;
}

// benchmark.cpp:293-317
void Benchmark::phase_1()
{
    phase_0_and_1_helper(1);
    std::cout << "\n";
    for (const std::string &a : settings_.algorithms)
        std::cout << "Optimal setup for algorithm " << a << ":\n\tData layout: " << optimal_data_layouts_[a]
                  << "\n\tWork-group size: " << optimal_work_group_sizes_[a]
                  << "\n\tAverage total time: " << optimal_total_times_[a] << "\n";
    if (settings_.test_linear_two_lattice)
        std::cout << "Optimal setup for algorithm lptl (Linear Pull Two-Lattice):\n\tData layout: "
                  << optimal_data_layouts_["lptl"] << "\n\tAverage total time: " << optimal_total_times_["lptl"]
                  << "\n";
    std::cout << "\n";
}

// benchmark.cpp:351-383
void Benchmark::phase_2_inner_loop(const std::string &ofdir, JsonValue &file_data, Progress &progress)
{
    for (unsigned int t = 0; t < settings_.test_runs; ++t)
    {
        benchmark_timer_.restart();
        const Run r = run_once();
        push_run(file_data, r);
        progress.advance(benchmark_timer_.elapsed());
        print_progress(progress);
        write(ofdir, file_data);
    }
}

// benchmark.cpp:385-449
void Benchmark::phase_2()
{
    std::cout << std::string(100, '=') << "\nStarting benchmark phase 2.\n\n";
    unsigned int total =
        static_cast<unsigned int>(settings_.algorithms.size() * settings_.phase_2_scenarios.size() * settings_.test_runs);
    if (settings_.test_linear_two_lattice)
        total += static_cast<unsigned int>(settings_.phase_2_scenarios.size() * settings_.test_runs);
    Progress progress(total);
    std::vector<std::string> algorithms = settings_.algorithms;
    if (settings_.test_linear_two_lattice)
        algorithms.push_back("lptl");
    for (const std::string &algorithm : algorithms)
    {
        properties_->algorithm = algorithm;
        properties_->data_layout = optimal_data_layouts_[algorithm];
        // a phase-1 winner, or (phase 1 not run) the lptl default and 256 otherwise
        const unsigned int wg = optimal_work_group_sizes_[algorithm];
        properties_->work_group_size = algorithm == "lptl" || wg <= 1 ? 256u : wg;
        for (const std::string &scenario : settings_.phase_2_scenarios)
        {
            properties_->scenario = scenario;
            file_interaction::properties_to_json(*properties_);
            const std::string ofdir = (project_root_ / "benchmarks" / "phase2").string() + "/phase2_" + algorithm +
                                      "_" + scenario + "_" + device_name_ + ".json";
            JsonValue d = file_header(algorithm);
            d["dataLayout"] = JsonValue::string(properties_->data_layout);
            d["workGroupSize"] = algorithm == "lptl" ? JsonValue::string("auto")
                                                     : JsonValue::unsigned_number(properties_->work_group_size);
            d["scenario"] = JsonValue::string(scenario);
            JsonValue file_data = run_arrays(d);
            phase_2_inner_loop(ofdir, file_data, progress);
        }
    }
    std::cout << "\nSuccessfully completed phase 2 of benchmark\n" << std::string(100, '=') << "\n\n";
}

}  // namespace benchmark
}  // namespace lbm
