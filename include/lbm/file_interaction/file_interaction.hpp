// Settings file <-> Properties.  Same functions, keys and semantics as the
// reference (include/lbm/file_interaction/file_interaction.hpp:30-49,
// src/lbm/file_interaction/file_interaction.cpp:11-69):
//   algorithmic.{algorithm, dataLayout, debugMode, frameUpdateInterval,
//                timeSteps, workGroupSize}
//   domain.{horizontalNodes, verticalNodes, scenario}
//   physical.{inletDensity, inletVelocity.{x,y}, outletDensity,
//             outletVelocity.{x,y}, relaxationTime}
// plus the optional section "mi355x" holding Properties::extensions.
#ifndef LBM_FILE_INTERACTION_HPP
#define LBM_FILE_INTERACTION_HPP

#include <string>

#include "../core/properties.hpp"
#include "json.hpp"

namespace lbm
{
namespace file_interaction
{

// Writes `properties` to properties.settings_path.  Node counts are written
// as stored (ghost layers included), as the reference does.
void properties_to_json(const core::Properties &properties);

// Reads a settings file; `offset` is added to both node counts before the
// ghost layers are added.  Records `path` in settings_path.
core::Properties json_to_properties(const std::string &path, int offset = 0);

// The same from an in-memory document.
core::Properties json_value_to_properties(const JsonValue &document, int offset = 0);
JsonValue properties_to_json_value(const core::Properties &properties);

}  // namespace file_interaction
}  // namespace lbm

#endif  // LBM_FILE_INTERACTION_HPP
