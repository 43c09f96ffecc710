// Project-wide constants of the D2Q9 lattice Boltzmann library.
// Mirrors include/lbm/core/constants.hpp:14-59 of the reference (same names,
// same values) so that callers of the reference compile unchanged.
#ifndef LBM_CONSTANTS_HPP
#define LBM_CONSTANTS_HPP

#include <array>
#include <string_view>

namespace lbm
{
// The plain path computes in double precision.  Reduced-precision storage is
// a run-time choice of the GPU library (Properties::extensions.storage),
// not a different real_type.
using real_type = double;

namespace core
{
namespace constants
{
constexpr unsigned int dimension_count = 2;
constexpr unsigned int direction_count = 9;

// Directions are numbered row by row from the lower left: d = 3 (cy + 1) + (cx + 1).
constexpr std::array<unsigned int, 8> streaming_directions = {0, 1, 2, 3, 5, 6, 7, 8};
constexpr std::array<unsigned int, 9> all_directions = {0, 1, 2, 3, 4, 5, 6, 7, 8};

constexpr std::array<real_type, 9> weights = {1.0 / 36, 1.0 / 9,  1.0 / 36, 1.0 / 9, 4.0 / 9,
                                              1.0 / 9,  1.0 / 36, 1.0 / 9,  1.0 / 36};

constexpr real_type boltzmann_constant = 1.380649e-23;

constexpr std::array<std::string_view, 4> algorithms = {"nptl", "lptl", "npol", "nsol"};
constexpr std::array<std::string_view, 3> data_layouts = {"collision", "stream", "bundle"};
constexpr std::array<std::string_view, 8> scenarios = {"Hagen-Poiseuille", "walls",      "circle", "square",
                                                       "wing",             "skyscraper", "porous", "plate"};

// Scenarios this library adds: a seeded porous channel of non-overlapping
// discs and a single disc in a channel.
constexpr std::array<std::string_view, 2> extra_scenarios = {"porous-circles", "cylinder"};

// Population storage types of the GPU library.
constexpr std::array<std::string_view, 4> storage_types = {"f64", "f32", "f16-shifted", "bf16-shifted"};
}  // namespace constants
}  // namespace core
}  // namespace lbm

#endif  // LBM_CONSTANTS_HPP
