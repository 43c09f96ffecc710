// Expanded lattice extents per algorithm.  The extents are part of the public
// surface (memory layout of the lattice, work-group tiling of the kernels) and
// follow the reference exactly: include/lbm/core/domain.hpp:29-85,
// src/lbm/core/simulation.cpp:179-382, include/lbm/core/simulation.hpp:373-446.
#ifndef LBM_DOMAIN_HPP
#define LBM_DOMAIN_HPP

#include <cstddef>
#include <string>

#include "properties.hpp"

namespace lbm
{
namespace core
{

class Domain
{
  public:
    unsigned int total_node_count;            // including buffers and padding
    unsigned int horizontal_nodes;            // including buffers and padding
    unsigned int vertical_nodes;              // including buffers and padding
    unsigned int subdomain_vertical_nodes;
    unsigned int subdomain_horizontal_nodes;
    unsigned int subdomain_count_vertical;
    unsigned int subdomain_count_horizontal;

    // lptl keeps the unexpanded lattice; nptl pads to whole tiles; npol and
    // nsol add a buffer row/column between tiles.  nsol with a work-group
    // size below 6 rewrites the settings file (if any) and throws
    // PropertyArgumentException (simulation.cpp:336-353).
    explicit Domain(Properties &properties);

  private:
    Domain(unsigned int total_node_count, unsigned int vertical_nodes, unsigned int horizontal_nodes,
           unsigned int subdomain_vertical_nodes, unsigned int subdomain_horizontal_nodes,
           unsigned int subdomain_count_vertical, unsigned int subdomain_count_horizontal);

    void setup_for_decomposed_two_lattice(unsigned int unexpanded_horizontal_nodes,
                                          unsigned int unexpanded_vertical_nodes, std::size_t work_group_size);
    void setup_for_buffered_two_lattice(unsigned int unexpanded_horizontal_nodes,
                                        unsigned int unexpanded_vertical_nodes, std::size_t work_group_size);
    void setup_for_swap(unsigned int unexpanded_horizontal_nodes, unsigned int unexpanded_vertical_nodes,
                        std::size_t work_group_size, const std::string &settings_path);
};

// Tile shape for an even work-group size: 4^k -> 2^k x 2^k; 2^p (p odd) ->
// 2^(p/2) rows x 2^(p/2+1) columns; other even sizes -> 2 x (wg/2).
void power_of_two_handling(std::size_t work_group_size, unsigned int &subdomain_vertical_nodes,
                           unsigned int &subdomain_horizontal_nodes);

namespace power_functions
{
// true iff x = 4^k with k >= 1 (x must be non-zero)
bool is_power_of_4(std::size_t x);
// k if x = 4^k (k >= 1), else 0
std::size_t which_power_of_4(std::size_t x);
// p if x = 2^p (p >= 1), else 0
std::size_t which_power_of_2(std::size_t x);
}  // namespace power_functions

}  // namespace core
}  // namespace lbm

#endif  // LBM_DOMAIN_HPP
