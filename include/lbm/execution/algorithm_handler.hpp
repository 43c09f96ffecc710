// Backend-neutral interface of a running simulation.  Identical to the
// reference's execution::AlgorithmHandler
// (include/lbm/execution/algorithm_handler.hpp:40-156) so that its callers
// (drivers, GUI, benchmark, tests) work unchanged.
#ifndef LBM_ALGORITHM_HANDLER_HPP
#define LBM_ALGORITHM_HANDLER_HPP

#include <concepts>
#include <cstddef>
#include <vector>

#include "../core/constants.hpp"
#include "../core/domain.hpp"
#include "../core/properties.hpp"

namespace lbm
{
namespace execution
{

class AlgorithmHandler
{
  public:
    // Upper bound of work-items per work-group on the device.
    std::size_t processing_element_constraint;

    virtual ~AlgorithmHandler() = default;

    // Rebuild the simulation from the settings (discarding the current state).
    virtual void initialize() = 0;
    // Run the remaining time steps on a worker thread.
    virtual void start() = 0;
    // Stop after the step batch in flight.
    virtual void pause() = 0;
    virtual bool is_finished() const = 0;
    // Join the worker; throws exceptions::algorithm::WaitException when there
    // is nothing to wait for or the worker failed.
    virtual void block_until_finished() = 0;

    // Inner-grid fields ((H-2)*(V-2), row-major), as last copied to the host.
    // "Velocity" is the momentum sum f c, never divided by the density.
    virtual std::vector<real_type> &get_densities() const = 0;
    virtual std::vector<real_type> &get_x_velocities() const = 0;
    virtual std::vector<real_type> &get_y_velocities() const = 0;
    virtual std::vector<real_type> &get_absolute_velocities() const = 0;

    // Unexpanded extents including the ghost layers.
    virtual unsigned int get_horizontal_nodes() const = 0;
    virtual unsigned int get_vertical_nodes() const = 0;
    virtual real_type get_inlet_density() const = 0;
    virtual real_type get_outlet_density() const = 0;
    virtual real_type get_progress() const = 0;
    // Milliseconds per time step of the last step batch.
    virtual real_type get_last_frametime() const = 0;
    virtual unsigned int get_current_iteration() const = 0;
    virtual const core::Properties &get_properties() const = 0;
    virtual const core::Domain &get_domain() const = 0;

  protected:
    explicit AlgorithmHandler(std::size_t processing_element_constraint)
        : processing_element_constraint(processing_element_constraint)
    {
    }
};

template <class T>
concept AlgorithmHandlerConcept = std::derived_from<T, AlgorithmHandler>;

}  // namespace execution
}  // namespace lbm

#endif  // LBM_ALGORITHM_HANDLER_HPP
