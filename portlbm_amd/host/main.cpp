// Command-line driver (reference: src/main_gpu.cpp:88-99,151-202).
//   portlbm [settings.json]               runs the configured simulation to the end
//   portlbm --benchmark [settings.json]   the three-phase benchmark (the reference's
//                                         BENCHMARK_MODE build); results go to
//                                         <project>/benchmarks/phase{0,1,2}/, <project>
//                                         being the settings file's grandparent directory
#include <filesystem>
#include <iostream>
#include <string>

#include "lbm/benchmark/benchmark.hpp"
#include "lbm/execution/hip_algorithm_handler.hpp"
#include "lbm/portlbm.hpp"

namespace
{
std::filesystem::path project_root(const char *argv0)
{
    // the executable lives in <root>/portlbm_amd/bin/
    std::error_code ec;
    std::filesystem::path exe = std::filesystem::canonical("/proc/self/exe", ec);
    if (ec)
        exe = std::filesystem::absolute(argv0);
    return exe.parent_path().parent_path().parent_path();
}

std::string resolve_settings(int argc, const char *const *argv)
{
    if (argc > 1)
        return argv[1];
    const std::filesystem::path candidate = project_root(argv[0]) / "settings" / "settings.json";
    if (!std::filesystem::exists(candidate))
        throw std::runtime_error("Cannot locate settings/settings.json relative to the executable.\n"
                                 "Usage: portlbm [/path/to/settings.json]");
    return candidate.string();
}
}  // namespace

int run_benchmark(const std::string &settings_path)
{
    const std::filesystem::path root = std::filesystem::absolute(settings_path).parent_path().parent_path();
    for (const char *phase : {"phase0", "phase1", "phase2"})
        std::filesystem::create_directories(root / "benchmarks" / phase);
    std::cout << "Starting benchmark.\n\n";
    auto handler = std::make_shared<lbm::execution::HipAlgorithmHandler>(settings_path);
    lbm::benchmark::Benchmark benchmark(handler);
    benchmark.phase_0();
    benchmark.phase_1();
    benchmark.phase_2();
    std::cout << "\nDone with all benchmarks. Exiting.\n\n";
    return 0;
}

int main(int argc, char *argv[])
{
    try
    {
        if (argc > 1 && std::string(argv[1]) == "--benchmark")
            return run_benchmark(resolve_settings(argc - 1, argv + 1));
        const std::string settings_path = resolve_settings(argc, argv);
        auto handler = lbm::create_handler(settings_path);
        lbm::core::Timer timer;
        handler->start();
        handler->block_until_finished();
        const double seconds = timer.elapsed();
        const auto &p = handler->get_properties();
        const double updates = static_cast<double>(p.domain_node_count) * handler->get_current_iteration();
        std::cout << p.to_string() << "steps: " << handler->get_current_iteration() << "  wall: " << seconds
                  << " s  MLUPS (inner nodes): " << updates / seconds / 1e6 << "\n";
    }
    catch (const lbm::exceptions::Exception &e)
    {
        std::cerr << e.to_string() << "\n";
        return 1;
    }
    catch (const std::exception &e)
    {
        std::cerr << e.what() << "\n";
        return 1;
    }
    return 0;
}
