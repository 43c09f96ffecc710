// AlgorithmHandler on the MI355X GPU library (include/lbm_hip.h).  Replaces
// the reference's SYCLAlgorithmHandler + SYCLAlgorithm + NPTL/LPTL/NPOL/NSOL
// drivers (include/lbm/execution/sycl_algorithm_handler.hpp:45-455,
// sycl_algorithm.hpp:31-116, gpu/*/*.hpp): the same construction checks,
// lifecycle and accessors; the step loop runs batches of fused steps on a
// worker thread and copies the macroscopic fields to the host on frame steps
// (iteration % frameUpdateInterval == frameUpdateInterval - 1) and at the end.
#ifndef LBM_HIP_ALGORITHM_HANDLER_HPP
#define LBM_HIP_ALGORITHM_HANDLER_HPP

#include <atomic>
#include <cstdint>
#include <future>
#include <memory>
#include <optional>
#include <string>
#include <vector>

#include "../core/control.hpp"
#include "../core/domain.hpp"
#include "../core/properties.hpp"
#include "algorithm_handler.hpp"

struct lbm_ctx;

namespace lbm
{
namespace execution
{

class HipAlgorithmHandler : public AlgorithmHandler
{
  public:
    // Reads the settings file (sycl_algorithm_handler.hpp:334-346).
    explicit HipAlgorithmHandler(const std::string &settings_path);
    // Uses the given properties (sycl_algorithm_handler.hpp:355-368).
    explicit HipAlgorithmHandler(core::Properties props);
    ~HipAlgorithmHandler() override;

    HipAlgorithmHandler(const HipAlgorithmHandler &) = delete;
    HipAlgorithmHandler &operator=(const HipAlgorithmHandler &) = delete;

    void initialize() override;
    void start() override;
    void pause() override;
    bool is_finished() const override;
    void block_until_finished() override;

    std::vector<real_type> &get_densities() const override { return densities_; }
    std::vector<real_type> &get_x_velocities() const override { return x_velocities_; }
    std::vector<real_type> &get_y_velocities() const override { return y_velocities_; }
    std::vector<real_type> &get_absolute_velocities() const override { return absolute_velocities_; }

    unsigned int get_horizontal_nodes() const override { return properties_->horizontal_nodes; }
    unsigned int get_vertical_nodes() const override { return properties_->vertical_nodes; }
    real_type get_inlet_density() const override { return properties_->inlet_density; }
    real_type get_outlet_density() const override { return properties_->outlet_density; }
    real_type get_progress() const override { return control_->get_progress(); }
    real_type get_last_frametime() const override { return control_->get_last_frametime(); }
    unsigned int get_current_iteration() const override { return control_->get_current_iteration(); }
    const core::Properties &get_properties() const override { return *properties_; }
    const core::Domain &get_domain() const override { return *domain_; }
    const std::string &get_settings_path() const { return settings_path_; }

    // Beyond the reference interface: the node-type table handed to the GPU,
    // the current populations (9 * expanded node count, layout order) and the
    // underlying C-ABI context.  Not to be used while the worker runs.
    const std::vector<std::int8_t> &get_phase() const { return phase_; }
    std::vector<double> get_populations() const;
    lbm_ctx *context() const { return ctx_; }
    // fluid nodes of the node-type table (the unit of MFLUPS)
    std::size_t fluid_node_count() const { return fluid_nodes_; }
    // GPU time of the steps run since the last initialize() (HIP events), seconds
    double device_run_seconds() const { return device_ms_ * 1e-3; }
    // name of the GPU the handler runs on
    static std::string device_name();

    // Snapshot file of the current state (settings + lattice state of the C
    // ABI's lbm_get_state): "PLBMSNAP", u32 version 1, u32 header length, a
    // JSON header {"settings", "iteration", "stateBytes"}, the state.  Loading
    // requires equal settings apart from timeSteps and frameUpdateInterval
    // (PropertyArgumentException otherwise) and continues bit for bit.  Not
    // while the worker runs.
    void save_snapshot(const std::string &path) const;
    void load_snapshot(const std::string &path);
    // The four result fields as CSV: header "x,y,density,x_velocity,y_velocity,
    // absolute_velocity", one row per inner node (x, y from 1, row-major).
    void export_csv(const std::string &path) const;

    // Debug mode (settings debugMode): the fields of every finished iteration,
    // appended in order (the reference's all_densities / all_x_velocities /
    // all_y_velocities, nptl.hpp:170-175), each (H-2)(V-2) values per step.
    const std::vector<real_type> &debug_densities() const { return debug_densities_; }
    const std::vector<real_type> &debug_x_velocities() const { return debug_x_velocities_; }
    const std::vector<real_type> &debug_y_velocities() const { return debug_y_velocities_; }

    // Seconds the last initialize() took (the reference's BENCHMARK_MODE
    // member, sycl_algorithm_handler.hpp:252-271,281).
    double initialization_time = 0.0;

  private:
    void initialize_algorithm();
    void release();
    void execute();        // worker body
    void execute_debug();  // worker body in debug mode: split steps
    static void pass_callback(void *user, int pass, std::uint64_t iteration);
    void print_pass(int pass, std::uint64_t iteration);
    void copy_macroscopic_observables_to_cpu();

    std::string settings_path_;
    std::optional<core::Properties> explicit_props_;
    std::unique_ptr<core::Properties> properties_;
    std::unique_ptr<core::Domain> domain_;
    std::unique_ptr<core::Control> control_;
    std::vector<std::int8_t> phase_;
    mutable std::vector<real_type> densities_, x_velocities_, y_velocities_, absolute_velocities_;
    std::vector<real_type> debug_densities_, debug_x_velocities_, debug_y_velocities_;
    lbm_ctx *ctx_ = nullptr;
    std::size_t fluid_nodes_ = 0;
    double device_ms_ = 0.0;
    std::future<void> future_;
    bool active_ = false;
};

}  // namespace execution
}  // namespace lbm

#endif  // LBM_HIP_ALGORITHM_HANDLER_HPP
