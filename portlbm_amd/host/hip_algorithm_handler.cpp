// HipAlgorithmHandler: the reference's handler lifecycle on the GPU library.
#include "lbm/execution/hip_algorithm_handler.hpp"
#include "lbm/core/timer.hpp"

#include <cstring>
#include <fstream>
#include <cstdio>
#include <algorithm>
#include <iostream>
#include <string>

#include "lbm/core/access.hpp"
#include "lbm/core/domain_initialization.hpp"
#include "lbm/exceptions/exceptions.hpp"
#include "lbm/file_interaction/file_interaction.hpp"
#include "lbm/portlbm_factory.hpp"
#include "lbm_hip.h"

namespace lbm
{
namespace execution
{

namespace
{
// Steps per lbm_run_frames call: at most max_batch, and about
// batch_target_ms of device time once a batch has been timed -- long enough to
// keep synchronisation cost negligible, short enough for pause() to take
// effect quickly on large lattices (0.4 ms per step at 8192 x 2048 in FP64).
constexpr unsigned int max_batch = 512;
constexpr double batch_target_ms = 20.0;

void check(int rc, const char *what)
{
    if (rc != LBM_OK)
    {
        const std::string msg = std::string(what) + ": " + lbm_last_error();
        if (rc == LBM_ERR_INVALID_ARGUMENT)
            throw exceptions::json::PropertyArgumentException(msg);
        throw exceptions::Exception(msg);
    }
}

int algorithm_id(const std::string &a)
{
    if (a == "nptl")
        return LBM_ALG_NPTL;
    if (a == "lptl")
        return LBM_ALG_LPTL;
    if (a == "npol")
        return LBM_ALG_NPOL;
    if (a == "nsol")
        return LBM_ALG_NSOL;
    throw exceptions::Exception("Unknown algorithm: " + a);
}

int layout_id(const std::string &l)
{
    if (l == "collision")
        return LBM_LAYOUT_COLLISION;
    if (l == "stream")
        return LBM_LAYOUT_STREAM;
    if (l == "bundle")
        return LBM_LAYOUT_BUNDLE;
    throw exceptions::Exception("Unknown data layout: " + l);
}

int storage_id(const std::string &s)
{
    if (s == "f64")
        return LBM_STORE_F64;
    if (s == "f32")
        return LBM_STORE_F32;
    if (s == "f16-shifted")
        return LBM_STORE_F16_SHIFTED;
    if (s == "bf16-shifted")
        return LBM_STORE_BF16_SHIFTED;
    throw exceptions::json::PropertyArgumentException("Unknown storage: " + s);
}
}  // namespace

HipAlgorithmHandler::HipAlgorithmHandler(const std::string &settings_path)
    : AlgorithmHandler(0), settings_path_(settings_path)
{
    initialize_algorithm();
}

HipAlgorithmHandler::HipAlgorithmHandler(core::Properties props)
    : AlgorithmHandler(0), settings_path_(props.settings_path), explicit_props_(std::move(props))
{
    initialize_algorithm();
}

HipAlgorithmHandler::~HipAlgorithmHandler() { release(); }

void HipAlgorithmHandler::release()
{
    if (control_)
        control_->forbid_execution();
    if (future_.valid())
        future_.wait();
    if (ctx_)
        lbm_destroy(ctx_);
    ctx_ = nullptr;
    active_ = false;
}

void HipAlgorithmHandler::initialize_algorithm()
{
    core::Timer timer;
    core::Properties props =
        explicit_props_.has_value() ? *explicit_props_ : file_interaction::json_to_properties(settings_path_);
    const int alg = algorithm_id(props.algorithm);
    const int layout = layout_id(props.data_layout);
    const int storage = storage_id(props.extensions.storage);

    uint32_t device_max = 0;
    check(lbm_device_max_work_group_size(&device_max), "device query");
    processing_element_constraint = device_max;
    // simulation.cpp:487-499
    if (props.work_group_size > device_max)
        throw exceptions::json::PropertyArgumentException(
            "work_group_size " + std::to_string(props.work_group_size) + " exceeds the device maximum of " +
            std::to_string(device_max) + "; reduce it to " + std::to_string(device_max) + " and reconstruct.");

    properties_ = std::make_unique<core::Properties>(props);
    control_ = std::make_unique<core::Control>(properties_->time_steps);
    domain_ = std::make_unique<core::Domain>(*properties_);

    // Results start at density -1 and zero momentum (simulation.cpp:386-417).
    const std::size_t D = properties_->domain_node_count;
    densities_.assign(D, -1.0);
    x_velocities_.assign(D, 0.0);
    y_velocities_.assign(D, 0.0);
    absolute_velocities_.assign(D, 0.0);

    phase_ = core::domain_initialization::build_phase(*properties_, *domain_);

    lbm_desc desc{};
    desc.algorithm = alg;
    desc.layout = layout;
    desc.storage = storage;
    desc.nan_protection = properties_->extensions.nan_protection ? 1 : 0;
    desc.rounding = properties_->extensions.rounding == "stochastic" ? LBM_ROUND_STOCHASTIC : LBM_ROUND_NEAREST;
    desc.tile_kernel = properties_->extensions.tile_kernel == "sweep"    ? LBM_TILE_SWEEP
                       : properties_->extensions.tile_kernel == "block"  ? LBM_TILE_BLOCK
                       : properties_->extensions.tile_kernel == "staged" ? LBM_TILE_STAGED
                                                                         : LBM_TILE_AUTO;
    desc.horizontal_nodes = properties_->horizontal_nodes;
    desc.vertical_nodes = properties_->vertical_nodes;
    desc.expanded_horizontal_nodes = domain_->horizontal_nodes;
    desc.expanded_vertical_nodes = domain_->vertical_nodes;
    desc.subdomain_horizontal_nodes = domain_->subdomain_horizontal_nodes;
    desc.subdomain_vertical_nodes = domain_->subdomain_vertical_nodes;
    desc.subdomain_count_horizontal = domain_->subdomain_count_horizontal;
    desc.subdomain_count_vertical = domain_->subdomain_count_vertical;
    desc.work_group_size = properties_->work_group_size;
    desc.relaxation_time = properties_->relaxation_time;
    desc.inlet_density = properties_->inlet_density;
    desc.inlet_velocity_x = properties_->inlet_velocity_x;
    desc.inlet_velocity_y = properties_->inlet_velocity_y;
    desc.outlet_density = properties_->outlet_density;
    desc.outlet_velocity_x = properties_->outlet_velocity_x;
    desc.outlet_velocity_y = properties_->outlet_velocity_y;

    check(lbm_create(&desc, &ctx_), "lbm_create");
    check(lbm_set_phase(ctx_, phase_.data(), phase_.size()), "lbm_set_phase");
    check(lbm_init_populations(ctx_), "lbm_init_populations");
    check(lbm_sync(ctx_), "lbm_sync");
    fluid_nodes_ = static_cast<std::size_t>(std::count(phase_.begin(), phase_.end(), std::int8_t{0}));
    device_ms_ = 0.0;
    debug_densities_.clear();
    debug_x_velocities_.clear();
    debug_y_velocities_.clear();
    initialization_time = timer.elapsed();
    // a ready future, so that block_until_finished() before start() returns
    // at once, as in the reference (sycl_algorithm.hpp:50-53)
    future_ = std::async(std::launch::deferred, [] {});
    active_ = false;
}

void HipAlgorithmHandler::initialize()
{
    release();
    initialize_algorithm();
}

void HipAlgorithmHandler::copy_macroscopic_observables_to_cpu()
{
    check(lbm_get_macros(ctx_, densities_.data(), x_velocities_.data(), y_velocities_.data(),
                         absolute_velocities_.data(), densities_.size()),
          "lbm_get_macros");
}

// The reference's execute() loop (nptl.hpp:116-141) in batches of at most
// max_batch steps.  Frame steps inside a batch copy the results to pinned
// host memory on a second stream while the steps go on, as the reference's
// non-blocking queue->copy does (nptl.hpp:71-75); after the batch the latest
// of those frames goes to the host vectors.
void HipAlgorithmHandler::execute()
{
    const unsigned int fui = properties_->frame_update_interval;
    unsigned int size = 8;  // first batch: a short one to time a step
    while (control_->is_execution_allowed())
    {
        const unsigned int it = control_->get_current_iteration();
        const unsigned int remaining = control_->get_max_iterations() - it;
        const unsigned int batch = std::min(remaining, size);
        control_->reset_timer();
        check(lbm_run_frames(ctx_, batch, fui), "lbm_run_frames");
        double ms = 0.0;
        check(lbm_last_run_ms(ctx_, &ms), "lbm_last_run_ms");
        device_ms_ += ms;
        const double per_step = ms / batch;
        size = per_step > 0.0 ? static_cast<unsigned int>(std::clamp(batch_target_ms / per_step, 1.0,
                                                                      static_cast<double>(max_batch)))
                              : max_batch;
        control_->finalize_iterations(batch, control_->elapsed_ms());
        // first frame step at or after it: it + (fui - 1 - it % fui)
        if (fui && (fui - 1 - it % fui) < batch)
        {
            std::uint64_t frame = 0;
            check(lbm_get_frame(ctx_, densities_.data(), x_velocities_.data(), y_velocities_.data(),
                                absolute_velocities_.data(), densities_.size(), &frame),
                  "lbm_get_frame");
        }
    }
    if (control_->is_finished())
        copy_macroscopic_observables_to_cpu();
}

// Debug mode (nptl.hpp:165-514 and the lptl/npol/nsol debug drivers): one
// split step at a time -- stream, outlet, moments, collision as separate
// launches -- the lattice printed after every pass when it has at most
// kDebugPrintNodes nodes, and the fields of every iteration kept.
namespace
{
constexpr std::size_t kDebugPrintNodes = 400;
const char *const kPassNames[4] = {"after streaming", "after the outlet update", "after the moments",
                                   "after collision"};
}  // namespace

void HipAlgorithmHandler::pass_callback(void *user, int pass, std::uint64_t iteration)
{
    static_cast<HipAlgorithmHandler *>(user)->print_pass(pass, iteration);
}

void HipAlgorithmHandler::print_pass(int pass, std::uint64_t iteration)
{
    const std::size_t N = domain_->total_node_count;
    if (N > kDebugPrintNodes)
        return;
    std::vector<double> values(9 * N);
    if (pass == LBM_PASS_COLLIDE)
    {
        check(lbm_get_populations(ctx_, values.data(), values.size()), "lbm_get_populations");
        std::cout << "Lattice " << kPassNames[pass] << " (iteration " << iteration << "):\n";
    }
    else
    {
        check(lbm_get_pass_populations(ctx_, values.data(), values.size()), "lbm_get_pass_populations");
        std::cout << "Populations " << kPassNames[pass] << " (iteration " << iteration << "):\n";
    }
    const unsigned int hx = domain_->horizontal_nodes;
    char buf[32];
    for (std::size_t n = 0; n < N; ++n)
    {
        if (phase_[n] != 0)
            continue;
        std::cout << "  (" << n % hx << "," << n / hx << ")";
        for (int d = 0; d < 9; ++d)
        {
            // collision dumps are in the layout order of lbm_get_populations, pass dumps direction-major
            std::size_t i = d * N + n;
            if (pass == LBM_PASS_COLLIDE)
            {
                const auto node = static_cast<unsigned int>(n), dir = static_cast<unsigned int>(d);
                const auto total = static_cast<unsigned int>(N);
                if (properties_->data_layout == "collision")
                    i = core::access::CollisionAccessor::at(node, dir, total);
                else if (properties_->data_layout == "bundle")
                    i = core::access::BundleAccessor::at(node, dir, total);
                else
                    i = core::access::StreamAccessor::at(node, dir, total);
            }
            std::snprintf(buf, sizeof buf, " %.6f", values[i]);
            std::cout << buf;
        }
        std::cout << "\n";
    }
    std::cout << std::flush;
}

void HipAlgorithmHandler::execute_debug()
{
    const std::size_t D = properties_->domain_node_count;
    while (control_->is_execution_allowed())
    {
        control_->reset_timer();
        const unsigned int it = control_->get_current_iteration();
        std::cout << "===== Iteration " << it << " =====\n";
        check(lbm_run_split(ctx_, 1, &HipAlgorithmHandler::pass_callback, this), "lbm_run_split");
        copy_macroscopic_observables_to_cpu();
        debug_densities_.insert(debug_densities_.end(), densities_.begin(), densities_.begin() + D);
        debug_x_velocities_.insert(debug_x_velocities_.end(), x_velocities_.begin(), x_velocities_.begin() + D);
        debug_y_velocities_.insert(debug_y_velocities_.end(), y_velocities_.begin(), y_velocities_.begin() + D);
        double ms = 0.0;
        check(lbm_last_run_ms(ctx_, &ms), "lbm_last_run_ms");
        device_ms_ += ms;
        control_->finalize_iterations(1, control_->elapsed_ms());
        std::cout << "Finished iteration " << it + 1 << " after " << control_->get_last_frametime()
                  << " milliseconds.\n";
    }
    if (control_->is_finished())
        std::cout << "All done, exiting simulation.\n" << std::flush;
}

void HipAlgorithmHandler::start()
{
    if (!active_)
    {
        control_->allow_execution();
        if (properties_->debug_mode)
            future_ = std::async(std::launch::async, [this] { execute_debug(); });
        else
            future_ = std::async(std::launch::async, [this] { execute(); });
        active_ = true;
    }
}

void HipAlgorithmHandler::pause()
{
    control_->forbid_execution();
    if (future_.valid())
        future_.wait();
    active_ = false;
}

bool HipAlgorithmHandler::is_finished() const { return control_->is_finished(); }

void HipAlgorithmHandler::block_until_finished()
{
    try
    {
        future_.get();
    }
    catch (const std::exception &e)
    {
        active_ = false;
        throw exceptions::algorithm::WaitException(
            std::string("A thread accessed a blocking statement causing it to wait for an algorithm that has not "
                        "been launched. (") +
            e.what() + ")");
    }
    check(lbm_sync(ctx_), "lbm_sync");
    active_ = false;
}

namespace
{
constexpr char kSnapshotMagic[8] = {'P', 'L', 'B', 'M', 'S', 'N', 'A', 'P'};
// 2: the 16-bit forms store f_d - w_d rho_lin(x) (1: f_d - w_d)
constexpr std::uint32_t kSnapshotVersion = 3;

// settings that must agree between a snapshot and the handler loading it
file_interaction::JsonValue comparable_settings(const core::Properties &p)
{
    file_interaction::JsonValue v = file_interaction::properties_to_json_value(p);
    v["algorithmic"]["timeSteps"] = file_interaction::JsonValue::unsigned_number(0);
    v["algorithmic"]["frameUpdateInterval"] = file_interaction::JsonValue::unsigned_number(0);
    return v;
}
}  // namespace

void HipAlgorithmHandler::save_snapshot(const std::string &path) const
{
    if (active_)
        throw exceptions::Exception("save_snapshot: pause or finish the run first");
    check(lbm_sync(ctx_), "lbm_sync");
    size_t bytes = 0;
    check(lbm_state_size(ctx_, &bytes), "lbm_state_size");
    std::vector<char> state(bytes);
    check(lbm_get_state(ctx_, state.data(), bytes), "lbm_get_state");
    file_interaction::JsonValue head = file_interaction::JsonValue::object();
    head["settings"] = file_interaction::properties_to_json_value(*properties_);
    head["iteration"] = file_interaction::JsonValue::unsigned_number(control_->get_current_iteration());
    head["stateBytes"] = file_interaction::JsonValue::unsigned_number(bytes);
    const std::string text = head.dump(0);
    std::ofstream out(path, std::ios::binary);
    if (!out)
        throw exceptions::Exception("cannot write snapshot \"" + path + "\"");
    const std::uint32_t version = kSnapshotVersion, len = static_cast<std::uint32_t>(text.size());
    out.write(kSnapshotMagic, 8);
    out.write(reinterpret_cast<const char *>(&version), 4);
    out.write(reinterpret_cast<const char *>(&len), 4);
    out.write(text.data(), static_cast<std::streamsize>(text.size()));
    out.write(state.data(), static_cast<std::streamsize>(state.size()));
    if (!out)
        throw exceptions::Exception("writing snapshot \"" + path + "\" failed");
}

void HipAlgorithmHandler::load_snapshot(const std::string &path)
{
    if (active_)
        throw exceptions::Exception("load_snapshot: pause or finish the run first");
    std::ifstream in(path, std::ios::binary);
    if (!in)
        throw exceptions::Exception("cannot read snapshot \"" + path + "\"");
    char magic[8];
    std::uint32_t version = 0, len = 0;
    in.read(magic, 8);
    in.read(reinterpret_cast<char *>(&version), 4);
    in.read(reinterpret_cast<char *>(&len), 4);
    if (!in || std::memcmp(magic, kSnapshotMagic, 8) != 0 || version != kSnapshotVersion || len > (1u << 24))
        throw exceptions::Exception("\"" + path + "\" is not a snapshot of this library");
    std::string text(len, '\0');
    in.read(text.data(), len);
    const file_interaction::JsonValue head = file_interaction::JsonValue::parse(text);
    // node counts are stored with their ghost layers (properties_to_json_value)
    const core::Properties saved = file_interaction::json_value_to_properties(head.at("settings"), -2);
    if (comparable_settings(saved).dump(0) != comparable_settings(*properties_).dump(0))
        throw exceptions::json::PropertyArgumentException(
            "the snapshot was taken with other settings (only timeSteps and frameUpdateInterval may differ)");
    const std::uint64_t iteration = head.at("iteration").as_unsigned("iteration");
    if (iteration > control_->get_max_iterations())
        throw exceptions::json::PropertyArgumentException("the snapshot is past this run's timeSteps");
    size_t bytes = 0;
    check(lbm_state_size(ctx_, &bytes), "lbm_state_size");
    if (head.at("stateBytes").as_unsigned("stateBytes") != bytes)
        throw exceptions::json::PropertyArgumentException("the snapshot's state size does not match");
    std::vector<char> state(bytes);
    in.read(state.data(), static_cast<std::streamsize>(bytes));
    if (!in)
        throw exceptions::Exception("snapshot \"" + path + "\" is truncated");
    check(lbm_set_state(ctx_, state.data(), bytes), "lbm_set_state");
    control_->set_current_iteration(static_cast<unsigned int>(iteration));
    copy_macroscopic_observables_to_cpu();
}

void HipAlgorithmHandler::export_csv(const std::string &path) const
{
    std::ofstream out(path);
    if (!out)
        throw exceptions::Exception("cannot write \"" + path + "\"");
    const unsigned int w = properties_->horizontal_nodes - 2, h = properties_->vertical_nodes - 2;
    out << "x,y,density,x_velocity,y_velocity,absolute_velocity\n";
    char line[192];
    for (unsigned int y = 0; y < h; ++y)
        for (unsigned int x = 0; x < w; ++x)
        {
            const std::size_t i = x + static_cast<std::size_t>(y) * w;
            std::snprintf(line, sizeof line, "%u,%u,%.17g,%.17g,%.17g,%.17g\n", x + 1, y + 1, densities_[i],
                          x_velocities_[i], y_velocities_[i], absolute_velocities_[i]);
            out << line;
        }
    if (!out)
        throw exceptions::Exception("writing \"" + path + "\" failed");
}

std::string HipAlgorithmHandler::device_name()
{
    char buf[256];
    check(lbm_device_name(buf, sizeof buf), "lbm_device_name");
    return buf;
}

std::vector<double> HipAlgorithmHandler::get_populations() const
{
    std::vector<double> values(9ull * domain_->total_node_count);
    check(lbm_get_populations(ctx_, values.data(), values.size()), "lbm_get_populations");
    return values;
}

}  // namespace execution

std::unique_ptr<execution::AlgorithmHandler> create_handler(const core::Properties &props)
{
    return std::make_unique<execution::HipAlgorithmHandler>(props);
}

std::unique_ptr<execution::AlgorithmHandler> create_handler(const std::string &settings_path)
{
    return std::make_unique<execution::HipAlgorithmHandler>(settings_path);
}

namespace exceptions
{
std::string Exception::to_string() const
{
    return std::string(exception_name_) +
           " thrown:\n"
           "-------------------------------------------------------------------------------\n"
           "What:\t\t" +
           what() + "\nFile:\t\t" + source_location_.file_name() + "\nFunction:\t" +
           source_location_.function_name() + "\nLine:\t\t" + std::to_string(source_location_.line()) + "\n";
}
}  // namespace exceptions

}  // namespace lbm
