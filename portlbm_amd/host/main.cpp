// Command-line driver: portlbm [settings.json] (reference: src/main_gpu.cpp:88-99,186-202).
// Runs the configured simulation to the end and prints its throughput.
#include <filesystem>
#include <iostream>
#include <string>

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

int main(int argc, char *argv[])
{
    try
    {
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
