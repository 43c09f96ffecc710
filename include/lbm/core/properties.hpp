// Run parameters of a simulation.  Field names, meaning and the ghost-layer
// accounting follow the reference's core::Properties
// (include/lbm/core/properties.hpp:30-166, src/lbm/core/simulation.cpp:21-125).
#ifndef LBM_PROPERTIES_HPP
#define LBM_PROPERTIES_HPP

#include <cstdint>
#include <string>

#include "constants.hpp"

namespace lbm
{
namespace core
{

// Options of this library that the reference does not have.  They live in an
// optional "mi355x" section of the settings file, so a reference settings
// file reads and writes back unchanged.
struct Extensions
{
    // population storage on the GPU: "f64", "f32", "f16-shifted", "bf16-shifted"
    std::string storage = "f64";
    // rounding of the FP16/BF16 population stores: "nearest" or "stochastic"
    std::string rounding = "nearest";
    // npol/nsol step kernel: "auto", "sweep" (one wave per tile) or "block"
    // (one work-group per tile, buffer refresh launches)
    std::string tile_kernel = "auto";
    // outputs that are NaN or above FLT_MAX are written as 0 (the reference's
    // WITH_NAN_PROTECTION build option, ON by default)
    bool nan_protection = true;
    // "porous-circles": seeded disc placement
    std::uint64_t seed = 20261202ull;
    double porosity = 0.7;
    double radius_min = 4.0;
    double radius_max = 16.0;
    unsigned int obstacle_margin = 32;
    std::uint64_t max_attempts = 1ull << 24;  // disc candidates tried at most
    // "cylinder": disc centred at (cylinder_x, (V-1)/2)
    unsigned int cylinder_x = 1024;
    double cylinder_radius = 64.0;

    bool is_default() const;
};

struct Properties
{
    /* ---- algorithmic options ---- */
    std::string algorithm;    // "nptl", "lptl", "npol", "nsol"
    std::string data_layout;  // "collision", "stream", "bundle"
    bool debug_mode;
    unsigned int work_group_size;
    unsigned int time_steps;
    unsigned int frame_update_interval;

    /* ---- domain options ---- */
    std::string scenario;
    unsigned int vertical_nodes;               // inner rows + 2 ghost rows
    unsigned int horizontal_nodes;             // inner columns + 2 ghost columns
    unsigned int total_unexpanded_node_count;  // with ghosts, without buffers
    unsigned int domain_node_count;            // inner nodes only

    /* ---- physical parameters ---- */
    real_type inlet_velocity_x;
    real_type inlet_velocity_y;
    real_type inlet_density;
    real_type outlet_velocity_x;
    real_type outlet_velocity_y;
    real_type outlet_density;
    real_type relaxation_time;

    // Path of the settings file this object was read from ("" if built in code).
    std::string settings_path;

    Extensions extensions;

    // `vertical_nodes` / `horizontal_nodes` are the inner extents; two ghost
    // layers are added in each direction (simulation.cpp:45-48).
    explicit Properties(std::string algorithm, std::string data_layout, bool debug_mode,
                        unsigned int work_group_size, unsigned int time_steps, unsigned int frame_update_interval,
                        std::string scenario, unsigned int vertical_nodes, unsigned int horizontal_nodes,
                        real_type inlet_velocity_x, real_type inlet_velocity_y, real_type inlet_density,
                        real_type outlet_velocity_x, real_type outlet_velocity_y, real_type outlet_density,
                        real_type relaxation_time);

    // Throws exceptions::json::PropertyArgumentException on the first invalid
    // field (simulation.cpp:60-125).
    void validate() const;

    std::string to_string() const;
};

}  // namespace core
}  // namespace lbm

#endif  // LBM_PROPERTIES_HPP
