// Factory: the preferred entry point (reference: include/lbm/portlbm_factory.hpp:59,73,
// src/lbm/portlbm_factory.cpp:17-25).  Callers see only AlgorithmHandler; the
// GPU library stays behind it.
#ifndef LBM_PORTLBM_FACTORY_HPP
#define LBM_PORTLBM_FACTORY_HPP

#include <memory>
#include <string>

#include "execution/algorithm_handler.hpp"

namespace lbm
{
std::unique_ptr<execution::AlgorithmHandler> create_handler(const core::Properties &props);
std::unique_ptr<execution::AlgorithmHandler> create_handler(const std::string &settings_path);
}  // namespace lbm

#endif  // LBM_PORTLBM_FACTORY_HPP
