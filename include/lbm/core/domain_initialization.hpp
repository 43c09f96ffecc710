// Node-type (phase) tables of the scenarios and the host equilibrium.
//
// Phase values: -1 ghost, buffer or padding node; 0 fluid; 1 solid.  The
// stencils keep the reference's unsigned-integer semantics
// (include/lbm/core/domain_initialization.hpp:41-192): differences of
// unsigned coordinates wrap, so e.g. the circle test compares true squares,
// and the walls stencil can yield 2 on very narrow grids (neither fluid nor
// solid).  The tables are built on the host and handed to the GPU library.
#ifndef LBM_DOMAIN_INITIALIZATION_HPP
#define LBM_DOMAIN_INITIALIZATION_HPP

#include <cstdint>
#include <vector>

#include "domain.hpp"
#include "properties.hpp"

namespace lbm
{
namespace core
{

// Equilibrium of (density, velocity) with the reference's host expression
// w_d (rho + 3 c.u + 4.5 (c.u)^2 - 1.5 |u|^2) (simulation.hpp:333-350).
void maxwell_boltzmann_distribution(real_type x_velocity, real_type y_velocity, real_type density,
                                    real_type *target);

constexpr unsigned int invert_direction(unsigned int dir) { return 8 - dir; }

namespace domain_initialization
{

namespace stencils
{
// Each returns the phase of inner node (x, y) for unexpanded extents V x H
// incl. ghosts (domain_initialization.hpp:43-191).
unsigned int hagen_poiseuille(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int walls(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int circle(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int square(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int plate(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int skyscraper(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
unsigned int wing(unsigned int x, unsigned int y, unsigned int V, unsigned int H);
}  // namespace stencils

struct Disc
{
    double x, y, r;
};

// Seeded placement of non-overlapping discs for the "porous-circles"
// scenario.  Candidates: r ~ U[radius_min, radius_max), centre
// x ~ U{x_lo..x_hi} with x_lo = 1 + margin, x_hi = H-2-margin, y ~ U{2..V-3},
// drawn from splitmix64(seed) as 53-bit uniforms in that order.  A candidate
// is rejected when it leaves [x_lo, x_hi] x [3, V-4] or comes within one node
// of an accepted disc (centre distance^2 < (r1 + r2 + 1)^2).  Placement stops
// when the solid node count reaches (1 - porosity) (H-2)(V-4) or after
// max_attempts candidates.  Node (x, y) is solid iff (x-cx)^2 + (y-cy)^2 <= r^2.
std::vector<Disc> porous_discs(unsigned int H, unsigned int V, const Extensions &ext);

// Phase table over the expanded lattice of `domain`, for any scenario of
// constants::scenarios or constants::extra_scenarios.  The "porous" scenario
// draws from the process-wide std::rand() state, as the reference does
// (domain_initialization.hpp:265-297), so two tables in one process differ.
std::vector<std::int8_t> build_phase(const Properties &properties, const Domain &domain);

}  // namespace domain_initialization
}  // namespace core
}  // namespace lbm

#endif  // LBM_DOMAIN_INITIALIZATION_HPP
