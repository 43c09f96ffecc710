// Properties, Domain extents, power functions, host equilibrium and the
// scenario node-type tables.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <sstream>
#include <string_view>
#include <vector>

#include "lbm/core/access.hpp"
#include "lbm/core/domain.hpp"
#include "lbm/core/domain_initialization.hpp"
#include "lbm/core/properties.hpp"
#include "lbm/exceptions/exceptions.hpp"
#include "lbm/file_interaction/file_interaction.hpp"

namespace lbm
{
namespace core
{

// ---------------------------------------------------------------------------
// Properties (src/lbm/core/simulation.cpp:21-175)
// ---------------------------------------------------------------------------
bool Extensions::is_default() const
{
    const Extensions d;
    return storage == d.storage && rounding == d.rounding && tile_kernel == d.tile_kernel &&
           nan_protection == d.nan_protection && seed == d.seed && porosity == d.porosity &&
           radius_min == d.radius_min && radius_max == d.radius_max && obstacle_margin == d.obstacle_margin &&
           max_attempts == d.max_attempts &&
           cylinder_x == d.cylinder_x && cylinder_radius == d.cylinder_radius;
}

Properties::Properties(std::string algorithm_, std::string data_layout_, bool debug_mode_,
                       unsigned int work_group_size_, unsigned int time_steps_, unsigned int frame_update_interval_,
                       std::string scenario_, unsigned int vertical_nodes_, unsigned int horizontal_nodes_,
                       real_type inlet_velocity_x_, real_type inlet_velocity_y_, real_type inlet_density_,
                       real_type outlet_velocity_x_, real_type outlet_velocity_y_, real_type outlet_density_,
                       real_type relaxation_time_)
    : algorithm(std::move(algorithm_)),
      data_layout(std::move(data_layout_)),
      debug_mode(debug_mode_),
      work_group_size(work_group_size_),
      time_steps(time_steps_),
      frame_update_interval(frame_update_interval_),
      scenario(std::move(scenario_)),
      vertical_nodes(vertical_nodes_ + 2),
      horizontal_nodes(horizontal_nodes_ + 2),
      total_unexpanded_node_count((vertical_nodes_ + 2) * (horizontal_nodes_ + 2)),
      domain_node_count(vertical_nodes_ * horizontal_nodes_),
      inlet_velocity_x(inlet_velocity_x_),
      inlet_velocity_y(inlet_velocity_y_),
      inlet_density(inlet_density_),
      outlet_velocity_x(outlet_velocity_x_),
      outlet_velocity_y(outlet_velocity_y_),
      outlet_density(outlet_density_),
      relaxation_time(relaxation_time_),
      settings_path("")
{
}

namespace
{
template <std::size_t K>
bool in_list(const std::string &value, const std::array<std::string_view, K> &list)
{
    return std::find(list.begin(), list.end(), std::string_view{value}) != list.end();
}
}  // namespace

void Properties::validate() const
{
    using exceptions::json::PropertyArgumentException;
    // stored node counts include the two ghost layers
    if (vertical_nodes < 4)
        throw PropertyArgumentException("vertical_nodes must be >= 4 (inner domain >= 2 rows)");
    if (horizontal_nodes < 4)
        throw PropertyArgumentException("horizontal_nodes must be >= 4 (inner domain >= 2 cols)");
    if (time_steps < 1)
        throw PropertyArgumentException("time_steps must be >= 1");
    // frame_update_interval > time_steps is allowed: results are then read at the end only
    if (frame_update_interval < 1)
        throw PropertyArgumentException("frame_update_interval must be >= 1");
    if (work_group_size < 1)
        throw PropertyArgumentException("work_group_size must be >= 1");
    if (inlet_density <= real_type{0})
        throw PropertyArgumentException("inlet_density must be > 0");
    if (outlet_density <= real_type{0})
        throw PropertyArgumentException("outlet_density must be > 0");
    if (relaxation_time <= real_type{0})
        throw PropertyArgumentException("relaxation_time must be > 0");
    if (!in_list(algorithm, constants::algorithms))
        throw PropertyArgumentException("Unknown algorithm: \"" + algorithm +
                                        "\". Valid values: lptl, nptl, npol, nsol");
    if (!in_list(data_layout, constants::data_layouts))
        throw PropertyArgumentException("Unknown data_layout: \"" + data_layout +
                                        "\". Valid values: stream, collision, bundle");
    if (!in_list(scenario, constants::scenarios) && !in_list(scenario, constants::extra_scenarios))
        throw PropertyArgumentException("Unknown scenario: \"" + scenario +
                                        "\". Valid values: Hagen-Poiseuille, walls, circle, square, plate, "
                                        "skyscraper, wing, porous, porous-circles, cylinder");
    if (!in_list(extensions.storage, constants::storage_types))
        throw PropertyArgumentException("Unknown storage: \"" + extensions.storage +
                                        "\". Valid values: f64, f32, f16-shifted, bf16-shifted");
    if (extensions.rounding != "nearest" && extensions.rounding != "stochastic")
        throw PropertyArgumentException("Unknown rounding: \"" + extensions.rounding +
                                        "\". Valid values: nearest, stochastic");
    if (extensions.tile_kernel != "auto" && extensions.tile_kernel != "sweep" && extensions.tile_kernel != "block" &&
        extensions.tile_kernel != "staged")
        throw PropertyArgumentException("Unknown tile kernel: \"" + extensions.tile_kernel +
                                        "\". Valid values: auto, sweep, block, staged");
    if (!(extensions.porosity > 0.0 && extensions.porosity <= 1.0))
        throw PropertyArgumentException("porosity must be in (0, 1]");
    if (!(extensions.radius_min > 0.0 && extensions.radius_max >= extensions.radius_min))
        throw PropertyArgumentException("obstacle radii must satisfy 0 < radiusMin <= radiusMax");
    if (!(extensions.cylinder_radius >= 0.0))
        throw PropertyArgumentException("cylinderRadius must be >= 0");
}

std::string Properties::to_string() const
{
    std::ostringstream os;
    os.setf(std::ios::fixed);
    os.precision(6);
    os << "Algorithmic properties: \n"
       << "\tAlgorithm: " << algorithm << " \n"
       << "\tData layout: " << data_layout << " \n"
       << "\tWork group size: " << work_group_size << " \n"
       << "\tDebug mode: " << (debug_mode ? "true" : "false") << " \n"
       << "\tTime steps: " << time_steps << " \n"
       << "\tFrame update interval: " << frame_update_interval << " \n"
       << "Domain properties: \n"
       << "\tScenario: " << scenario << "\n"
       << "\tVertical nodes: " << vertical_nodes << " \n"
       << "\tHorizontal nodes: " << horizontal_nodes << " \n"
       << "\tNode count: " << total_unexpanded_node_count << " \n"
       << "Physical: \n"
       << "\tInlet velocity: (" << inlet_velocity_x << ", " << inlet_velocity_y << ")\n"
       << "\tInlet density: " << inlet_density << "\n"
       << "\tOutlet velocity: (" << outlet_velocity_x << ", " << outlet_velocity_y << ")\n"
       << "\tOutlet density: " << outlet_density << "\n"
       << "\tRelaxation time: " << relaxation_time << "\n"
       << "Storage: " << extensions.storage << " (rounding: " << extensions.rounding << ")\n\n";
    return os.str();
}

// ---------------------------------------------------------------------------
// power functions (simulation.hpp:380-444)
// ---------------------------------------------------------------------------
namespace power_functions
{
bool is_power_of_4(std::size_t x)
{
    if (x & 1u)
        return false;
    while (!(x & 3u))
        x >>= 2;
    return !(x & 2u);
}

std::size_t which_power_of_4(std::size_t x)
{
    if (x & 1u)
        return 0;
    std::size_t k = 0;
    while (!(x & 3u))
    {
        x >>= 2;
        ++k;
    }
    return (x & 2u) ? 0 : k;
}

std::size_t which_power_of_2(std::size_t x)
{
    if (x & 1u)
        return 0;
    std::size_t p = 0;
    while (!(x & 1u))
    {
        x >>= 1;
        ++p;
    }
    return (x >> 1) ? 0 : p;
}
}  // namespace power_functions

// ---------------------------------------------------------------------------
// Domain (simulation.cpp:179-382)
// ---------------------------------------------------------------------------
void power_of_two_handling(std::size_t work_group_size, unsigned int &sub_v, unsigned int &sub_h)
{
    const std::size_t p4 = power_functions::which_power_of_4(work_group_size);
    const std::size_t p2 = power_functions::which_power_of_2(work_group_size);
    if (p4)
    {
        sub_h = sub_v = 1u << p4;  // square tiles
    }
    else if (p2)
    {
        sub_v = 1u << (p2 / 2);  // twice as wide as high
        sub_h = 2u * sub_v;
    }
    else
    {
        sub_v = 2;  // stripes
        sub_h = static_cast<unsigned int>(work_group_size / 2);
    }
}

namespace
{
unsigned int tiles_needed(unsigned int inner, unsigned int tile) { return (inner + tile - 1) / tile; }
}  // namespace

Domain::Domain(unsigned int total_node_count_, unsigned int vertical_nodes_, unsigned int horizontal_nodes_,
               unsigned int subdomain_vertical_nodes_, unsigned int subdomain_horizontal_nodes_,
               unsigned int subdomain_count_vertical_, unsigned int subdomain_count_horizontal_)
    : total_node_count(total_node_count_),
      horizontal_nodes(horizontal_nodes_),
      vertical_nodes(vertical_nodes_),
      subdomain_vertical_nodes(subdomain_vertical_nodes_),
      subdomain_horizontal_nodes(subdomain_horizontal_nodes_),
      subdomain_count_vertical(subdomain_count_vertical_),
      subdomain_count_horizontal(subdomain_count_horizontal_)
{
}

Domain::Domain(Properties &p)
    : Domain(p.horizontal_nodes * p.vertical_nodes, p.vertical_nodes, p.horizontal_nodes, p.vertical_nodes,
             p.horizontal_nodes, 1, 1)
{
    if (p.algorithm == "nptl")
        setup_for_decomposed_two_lattice(p.horizontal_nodes, p.vertical_nodes, p.work_group_size);
    else if (p.algorithm == "npol")
        setup_for_buffered_two_lattice(p.horizontal_nodes, p.vertical_nodes, p.work_group_size);
    else if (p.algorithm == "nsol")
        setup_for_swap(p.horizontal_nodes, p.vertical_nodes, p.work_group_size, p.settings_path);
}

void Domain::setup_for_decomposed_two_lattice(unsigned int H, unsigned int V, std::size_t wg)
{
    if (wg & 1u)
    {
        subdomain_horizontal_nodes = static_cast<unsigned int>(wg);
        subdomain_vertical_nodes = 1;
    }
    else
        power_of_two_handling(wg, subdomain_vertical_nodes, subdomain_horizontal_nodes);
    subdomain_count_vertical = tiles_needed(V - 2, subdomain_vertical_nodes);
    subdomain_count_horizontal = tiles_needed(H - 2, subdomain_horizontal_nodes);
    vertical_nodes = subdomain_vertical_nodes * subdomain_count_vertical + 2;
    horizontal_nodes = subdomain_horizontal_nodes * subdomain_count_horizontal + 2;
    total_node_count = vertical_nodes * horizontal_nodes;
}

void Domain::setup_for_buffered_two_lattice(unsigned int H, unsigned int V, std::size_t wg)
{
    if (wg & 1u)
    {
        subdomain_horizontal_nodes = static_cast<unsigned int>(wg);
        subdomain_vertical_nodes = 1;
    }
    else
        power_of_two_handling(wg, subdomain_vertical_nodes, subdomain_horizontal_nodes);
    subdomain_count_vertical = tiles_needed(V - 2, subdomain_vertical_nodes);
    subdomain_count_horizontal = tiles_needed(H - 2, subdomain_horizontal_nodes);
    vertical_nodes = (subdomain_vertical_nodes + 1) * subdomain_count_vertical + 1;
    horizontal_nodes = (subdomain_horizontal_nodes + 1) * subdomain_count_horizontal + 1;
    total_node_count = vertical_nodes * horizontal_nodes;
}

void Domain::setup_for_swap(unsigned int H, unsigned int V, std::size_t wg, const std::string &settings_path)
{
    constexpr std::size_t minimum = 6;
    if (wg < minimum)
    {
        // Rewrite the settings file with the minimum so that a restart works,
        // then refuse (simulation.cpp:336-353).
        if (!settings_path.empty())
        {
            Properties corrected = file_interaction::json_to_properties(settings_path, -2);
            corrected.work_group_size = minimum;
            file_interaction::properties_to_json(corrected);
        }
        throw exceptions::json::PropertyArgumentException(
            "Detected illegal work group size of " + std::to_string(wg) +
            " that is below the minimum work-group size of 6 of the swap algorithm. JSON property "
            "\"workGroupSize\" has been set to 6 to enable graceful program restart. ");
    }
    if (wg & 1u)
        --wg;
    power_of_two_handling(wg, subdomain_vertical_nodes, subdomain_horizontal_nodes);
    // which_power_of_4 reports 1 for every wg = 4m with m = 1 (mod 4), e.g.
    // 20, which gives 2 x 2 tiles; without the halo columns such a tile has
    // width 0 (the reference divides by zero here)
    if (subdomain_vertical_nodes < 2 || subdomain_horizontal_nodes < 3)
        throw exceptions::json::PropertyArgumentException(
            "work group size " + std::to_string(wg) + " gives a swap tile of width 0; choose another size");
    subdomain_vertical_nodes -= 1;    // room for the halo row of the work-group
    subdomain_horizontal_nodes -= 2;  // room for the two halo columns
    subdomain_count_vertical = tiles_needed(V - 2, subdomain_vertical_nodes);
    vertical_nodes = (subdomain_vertical_nodes + 1) * subdomain_count_vertical + 1;
    subdomain_count_horizontal = tiles_needed(H - 2, subdomain_horizontal_nodes);
    horizontal_nodes = (subdomain_horizontal_nodes + 1) * subdomain_count_horizontal + 1;
    total_node_count = vertical_nodes * horizontal_nodes;
}

// ---------------------------------------------------------------------------
// equilibrium (simulation.hpp:333-350)
// ---------------------------------------------------------------------------
void maxwell_boltzmann_distribution(real_type ux, real_type uy, real_type rho, real_type *target)
{
    for (unsigned int d = 0; d < 9; ++d)
    {
        const int cx = static_cast<int>(d % 3) - 1, cy = static_cast<int>(d / 3) - 1;
        const real_type cu = cx * ux + cy * uy;
        target[d] = constants::weights[d] * (rho + 3 * cu + 9.0 / 2 * std::pow(cu, 2) - 3.0 / 2 * (ux * ux + uy * uy));
    }
}

// ---------------------------------------------------------------------------
// scenarios
// ---------------------------------------------------------------------------
namespace domain_initialization
{
namespace stencils
{
// All arithmetic is on unsigned int, as in the reference: a difference that
// would be negative wraps and the wrapped square is the true square mod 2^32.
unsigned int hagen_poiseuille(unsigned int, unsigned int, unsigned int, unsigned int) { return 0; }

unsigned int walls(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    // thin plates hanging from the bottom at H/6, H/2, 5H/6 and standing on
    // the top at H/3, 2H/3; the two terms are summed (:64-71)
    unsigned int c = (x == H / 6 || x == (3 * H) / 6 || x == (5 * H) / 6) && y > 1 && y < V / 2;
    c += (x == (2 * H) / 6 || x == (4 * H) / 6) && y < V - 2 && y > (V - 1) / 2;
    return c;
}

unsigned int circle(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    const unsigned int cx = 2 * H / 10, cy = V / 2, r = V / 5;
    const unsigned int dx = cx - x, dy = cy - y;
    return y >= cy - r && y <= cy + r && x >= cx - r && x <= cx + r && dx * dx + dy * dy <= r * r;
}

unsigned int square(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    const unsigned int cx = 2 * H / 10, cy = V / 2, h = V / 10;
    return y >= cy - h && y <= cy + h && x >= cx - h && x <= cx + h;
}

unsigned int plate(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    const unsigned int cx = H / 10, cy = V / 2, hh = 7 * V / 40, hw = hh / 60;
    return y >= cy - hh && y <= cy + hh && x >= cx - hw && x <= cx + hw;
}

unsigned int skyscraper(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    const unsigned int cx = 3 * H / 10;
    const bool base = y >= 1 && y <= V / 4 && x >= cx - V / 16 && x <= cx + V / 16;
    const bool middle = y <= 2 * V / 5 && x >= cx - V / 24 && x <= cx + V / 24;
    const bool antenna = y <= V / 2 && x >= cx - 1 && x <= cx + 1;
    return base || middle || antenna;
}

unsigned int wing(unsigned int x, unsigned int y, unsigned int V, unsigned int H)
{
    // int centre, unsigned differences (:178-189); the 1.4-power term uses
    // the wrapped difference, so it only matters right of the centre
    const int cx = static_cast<int>(H / 5), cy = static_cast<int>(V / 2), rmax = static_cast<int>(V / 5);
    const unsigned int dx = x - static_cast<unsigned int>(cx), dy = y - static_cast<unsigned int>(cy);
    const unsigned int sqx = dx * dx, sqy = dy * dy;
    const bool ellipse = 0.2 * static_cast<double>(sqx) + static_cast<double>(sqy) <= rmax;
    const bool tail = std::pow(static_cast<double>(dx), 1.4) + static_cast<double>(25u * sqy) <= 25 * rmax;
    return ellipse || tail;
}
}  // namespace stencils

namespace
{
using StencilFn = unsigned int (*)(unsigned int, unsigned int, unsigned int, unsigned int);

StencilFn stencil_for(const std::string &scenario)
{
    if (scenario == "Hagen-Poiseuille")
        return stencils::hagen_poiseuille;
    if (scenario == "walls")
        return stencils::walls;
    if (scenario == "circle")
        return stencils::circle;
    if (scenario == "square")
        return stencils::square;
    if (scenario == "plate")
        return stencils::plate;
    if (scenario == "skyscraper")
        return stencils::skyscraper;
    if (scenario == "wing")
        return stencils::wing;
    return nullptr;
}

// splitmix64 stream with 53-bit uniforms in [0, 1).
struct SplitMix64
{
    std::uint64_t state;
    std::uint64_t next()
    {
        std::uint64_t z = (state += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return static_cast<double>(next() >> 11) * (1.0 / 9007199254740992.0); }
};

std::size_t disc_node_count(const Disc &c)
{
    const long reach = static_cast<long>(std::ceil(c.r));
    std::size_t n = 0;
    for (long y = static_cast<long>(c.y) - reach; y <= static_cast<long>(c.y) + reach; ++y)
        for (long x = static_cast<long>(c.x) - reach; x <= static_cast<long>(c.x) + reach; ++x)
        {
            const double dx = static_cast<double>(x) - c.x, dy = static_cast<double>(y) - c.y;
            n += dx * dx + dy * dy <= c.r * c.r;
        }
    return n;
}
}  // namespace

std::vector<Disc> porous_discs(unsigned int H, unsigned int V, const Extensions &ext)
{
    std::vector<Disc> discs;
    SplitMix64 rng{ext.seed};
    const double x_lo = 1.0 + ext.obstacle_margin;
    const double x_hi = static_cast<double>(H - 2) - ext.obstacle_margin;
    if (x_hi < x_lo || V < 8)
        return discs;
    const double target = (1.0 - ext.porosity) * static_cast<double>(H - 2) * static_cast<double>(V - 4);

    // uniform cell grid for the overlap test: any two discs that could touch
    // lie in neighbouring cells when the cell is at least 2 r_max + 1 wide
    const double cell = 2.0 * ext.radius_max + 1.0;
    const std::size_t gw = static_cast<std::size_t>(std::ceil((x_hi + 2.0) / cell)) + 1;
    const std::size_t gh = static_cast<std::size_t>(std::ceil((static_cast<double>(V) + 2.0) / cell)) + 1;
    std::vector<std::vector<std::uint32_t>> grid(gw * gh);

    double solid = 0.0;
    for (std::uint64_t attempt = 0; attempt < ext.max_attempts && solid < target; ++attempt)
    {
        const double r = ext.radius_min + (ext.radius_max - ext.radius_min) * rng.uniform();
        const double x = x_lo + std::floor(rng.uniform() * (x_hi - x_lo + 1.0));
        const double y = 2.0 + std::floor(rng.uniform() * static_cast<double>(V - 4));
        if (x - r < x_lo || x + r > x_hi || y - r < 3.0 || y + r > static_cast<double>(V) - 4.0)
            continue;
        const std::size_t gx = static_cast<std::size_t>(x / cell), gy = static_cast<std::size_t>(y / cell);
        bool clear = true;
        for (std::size_t j = gy ? gy - 1 : 0; j <= std::min(gh - 1, gy + 1) && clear; ++j)
            for (std::size_t i = gx ? gx - 1 : 0; i <= std::min(gw - 1, gx + 1) && clear; ++i)
                for (std::uint32_t k : grid[j * gw + i])
                {
                    const double dx = x - discs[k].x, dy = y - discs[k].y, g = r + discs[k].r + 1.0;
                    if (dx * dx + dy * dy < g * g)
                    {
                        clear = false;
                        break;
                    }
                }
        if (!clear)
            continue;
        grid[gy * gw + gx].push_back(static_cast<std::uint32_t>(discs.size()));
        discs.push_back({x, y, r});
        solid += static_cast<double>(disc_node_count(discs.back()));
    }
    return discs;
}

std::vector<std::int8_t> build_phase(const Properties &p, const Domain &domain)
{
    const unsigned int H = p.horizontal_nodes, V = p.vertical_nodes;
    const bool buffered = p.algorithm == "npol" || p.algorithm == "nsol";
    const unsigned int sv = domain.subdomain_vertical_nodes, sh = domain.subdomain_horizontal_nodes;
    const unsigned int Hx = domain.horizontal_nodes;
    auto index = [&](unsigned int x, unsigned int y) {
        return buffered ? access::decomposed::BufferedNodeAccess::get_index(x, y, sv, sh, Hx)
                        : access::decomposed::NonBufferedNodeAccess::get_index(x, y, sv, sh, Hx);
    };
    // everything starts as ghost/buffer/padding (simulation.cpp:424)
    std::vector<std::int8_t> phase(domain.total_node_count, -1);

    if (p.scenario == "porous")
    {
        // random obstacles from the process-wide std::rand() stream, then the
        // two walls (domain_initialization.hpp:265-297)
        for (unsigned int y = 2; y + 2 < V; ++y)
            for (unsigned int x = 1; x + 1 < H; ++x)
            {
                const int one = 1 + (std::rand() % 10);
                const int two = 1 + (std::rand() % 10);
                const bool solid = static_cast<int>(x) % one == one - 1 && static_cast<int>(y) % two == two - 1;
                phase[index(x, y)] = solid ? 1 : 0;
            }
        for (unsigned int y = 1; y < V; y += V - 3)
            for (unsigned int x = 1; x + 1 < H; ++x)
            {
                const unsigned int n = index(x, y);
                if (n < phase.size())
                    phase[n] = 1;
            }
        return phase;
    }

    // channel walls on rows 1 and V-2 (domain_initialization.hpp:214-239)
    for (unsigned int x = 1; x + 1 < H; ++x)
    {
        phase[index(x, 1)] = 1;
        phase[index(x, V - 2)] = 1;
    }

    if (p.scenario == "porous-circles" || p.scenario == "cylinder")
    {
        for (unsigned int y = 2; y + 2 < V; ++y)
            for (unsigned int x = 1; x + 1 < H; ++x)
                phase[index(x, y)] = 0;
        std::vector<Disc> discs;
        if (p.scenario == "cylinder")
            discs.push_back({static_cast<double>(p.extensions.cylinder_x), static_cast<double>((V - 1) / 2),
                             p.extensions.cylinder_radius});
        else
            discs = porous_discs(H, V, p.extensions);
        for (const Disc &c : discs)
        {
            const double reach = std::ceil(c.r);
            const long x0 = static_cast<long>(std::max(1.0, c.x - reach));
            const long x1 = static_cast<long>(std::min(static_cast<double>(H - 2), c.x + reach));
            const long y0 = static_cast<long>(std::max(2.0, c.y - reach));
            const long y1 = static_cast<long>(std::min(static_cast<double>(V - 3), c.y + reach));
            for (long y = y0; y <= y1; ++y)
                for (long x = x0; x <= x1; ++x)
                {
                    const double dx = static_cast<double>(x) - c.x, dy = static_cast<double>(y) - c.y;
                    if (dx * dx + dy * dy <= c.r * c.r)
                        phase[index(static_cast<unsigned int>(x), static_cast<unsigned int>(y))] = 1;
                }
        }
        return phase;
    }

    const StencilFn stencil = stencil_for(p.scenario);
    if (!stencil)
        throw exceptions::json::PropertyArgumentException("Unknown scenario: \"" + p.scenario + "\"");
    // stencil over x in [1, H-2], y in [2, V-3] (domain_initialization.hpp:427-435)
    for (unsigned int y = 2; y + 2 < V; ++y)
        for (unsigned int x = 1; x + 1 < H; ++x)
            phase[index(x, y)] = static_cast<std::int8_t>(stencil(x, y, V, H));
    return phase;
}

}  // namespace domain_initialization
}  // namespace core
}  // namespace lbm
