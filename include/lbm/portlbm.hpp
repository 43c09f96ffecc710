// Umbrella header of the library (reference: include/lbm/portlbm.hpp:29-53).
#ifndef PORTLBM_HPP
#define PORTLBM_HPP

#define PORTLBM_VERSION_MAJOR 0
#define PORTLBM_VERSION_MINOR 1
#define PORTLBM_VERSION_PATCH 0

#include "core/access.hpp"
#include "core/constants.hpp"
#include "core/control.hpp"
#include "core/domain.hpp"
#include "core/domain_initialization.hpp"
#include "core/properties.hpp"
#include "core/timer.hpp"
#include "exceptions/exceptions.hpp"
#include "execution/algorithm_handler.hpp"
#include "execution/hip_algorithm_handler.hpp"
#include "file_interaction/file_interaction.hpp"
#include "portlbm_factory.hpp"

#endif  // PORTLBM_HPP
