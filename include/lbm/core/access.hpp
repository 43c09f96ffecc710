// Population layouts and node/neighbour/result index arithmetic.
//
// These formulas define the memory formats of the library and therefore must
// equal the reference's (include/lbm/core/access.hpp:41-302): a caller that
// addresses a lattice or a result array with the reference's functions has to
// land on the same element here.
#ifndef LBM_ACCESS_HPP
#define LBM_ACCESS_HPP

#include <concepts>
#include <string_view>

#include "constants.hpp"

namespace lbm
{
namespace core
{
namespace access
{

// Array of structures: the nine populations of a node are adjacent
// (reference access.hpp:41-61).
struct CollisionAccessor
{
    constexpr static std::string_view layout_string = "collision";
    static constexpr unsigned int at(unsigned int node, unsigned int direction, unsigned int /*total*/)
    {
        return 9u * node + direction;
    }
};

// Structure of arrays: one contiguous array per direction (access.hpp:66-86).
struct StreamAccessor
{
    constexpr static std::string_view layout_string = "stream";
    static constexpr unsigned int at(unsigned int node, unsigned int direction, unsigned int total)
    {
        return total * direction + node;
    }
};

// Three arrays, one per row of directions (equal cy); inside each the three
// populations of a node are adjacent (access.hpp:91-111).
struct BundleAccessor
{
    constexpr static std::string_view layout_string = "bundle";
    static constexpr unsigned int at(unsigned int node, unsigned int direction, unsigned int total)
    {
        return 3u * (direction / 3u) * total + (direction % 3u) + 3u * node;
    }
};

template <class T>
concept AccessorConcept =
    std::same_as<T, CollisionAccessor> || std::same_as<T, StreamAccessor> || std::same_as<T, BundleAccessor>;

// Node reached from `node_index` by one step along `direction` on a row-major
// lattice `horizontal_nodes` wide (access.hpp:142-152).
inline unsigned int get_neighbor(unsigned int node_index, unsigned int direction, unsigned int horizontal_nodes)
{
    const int dy = static_cast<int>(direction / 3) - 1;
    const int dx = static_cast<int>(direction % 3) - 1;
    return node_index + dy * horizontal_nodes + dx;
}

// Row-major node index, origin at the lower left (access.hpp:164-167).
inline unsigned int get_node_index(unsigned int x, unsigned int y, unsigned int horizontal_nodes)
{
    return x + y * horizontal_nodes;
}

// Result index of inner node (x, y) with a time-series offset (access.hpp:186-193).
inline unsigned int get_result_index(unsigned int x, unsigned int y, unsigned int horizontal_nodes_domain,
                                     unsigned int domain_node_count, unsigned int time_step)
{
    return (x - 1) + (y - 1) * (horizontal_nodes_domain - 2) + time_step * domain_node_count;
}

// Result index of inner node (x, y): ghost layer stripped (access.hpp:208-212).
inline unsigned int get_result_index(unsigned int x, unsigned int y, unsigned int horizontal_nodes_domain)
{
    return (x - 1) + (y - 1) * (horizontal_nodes_domain - 2);
}

// Result index of an inner node given by its linear index (access.hpp:225-229).
inline unsigned int get_result_index(unsigned int node_index, unsigned int horizontal_nodes)
{
    return ((node_index - horizontal_nodes) / horizontal_nodes) * (horizontal_nodes - 2) +
           (node_index - 1) % horizontal_nodes;
}

namespace decomposed
{
// Node index on a lattice without buffer rows/columns (access.hpp:244-269).
struct NonBufferedNodeAccess
{
    static unsigned int get_index(unsigned int x, unsigned int y, unsigned int /*sub_v*/, unsigned int /*sub_h*/,
                                  unsigned int extended_horizontal_nodes)
    {
        return x + y * extended_horizontal_nodes;
    }
};

// Node index on a lattice with one buffer column after every sub_h real
// columns and one buffer row after every sub_v real rows (access.hpp:276-302).
struct BufferedNodeAccess
{
    static unsigned int get_index(unsigned int x, unsigned int y, unsigned int sub_v, unsigned int sub_h,
                                  unsigned int extended_horizontal_nodes)
    {
        return (x + ((x - 1) / sub_h)) + (y + ((y - 1) / sub_v)) * extended_horizontal_nodes;
    }
};

template <class T>
concept NodeAccessor = std::same_as<T, NonBufferedNodeAccess> || std::same_as<T, BufferedNodeAccess>;
}  // namespace decomposed

}  // namespace access
}  // namespace core
}  // namespace lbm

#endif  // LBM_ACCESS_HPP
