// Step counter, stop flag and frame time of a running simulation.  Same
// interface as the reference's core::Control (include/lbm/core/simulation.hpp:46-135);
// the stop flag is atomic because pause() is called from another thread than
// the one running the steps.
#ifndef LBM_CONTROL_HPP
#define LBM_CONTROL_HPP

#include <atomic>
#include <limits>

#include "constants.hpp"
#include "timer.hpp"

namespace lbm
{
namespace core
{

class Control
{
  public:
    explicit Control(unsigned int max_iterations) : max_iterations_(max_iterations) {}

    void forbid_execution() { stopped_.store(true); }
    void allow_execution() { stopped_.store(false); }
    bool is_execution_allowed() const { return !stopped_.load() && current_iteration_.load() < max_iterations_; }
    bool is_finished() const { return current_iteration_.load() == max_iterations_; }
    bool is_paused() const { return stopped_.load(); }

    // One iteration done (simulation.hpp:107-112).
    void finalize_iteration() { finalize_iterations(1, timer_.elapsed() * 1e3); }

    // `count` iterations done in `elapsed_ms`; the frame time is their mean.
    void finalize_iterations(unsigned int count, double elapsed_ms)
    {
        const unsigned int it = current_iteration_.load() + count;
        current_iteration_.store(it);
        progress_.store(static_cast<double>(it) / static_cast<double>(max_iterations_));
        if (count)
            last_frametime_.store(elapsed_ms / count);
    }

    // Continue from `it` completed iterations (a restored snapshot).
    void set_current_iteration(unsigned int it)
    {
        current_iteration_.store(it);
        progress_.store(static_cast<double>(it) / static_cast<double>(max_iterations_));
    }

    void reset_timer() { timer_.restart(); }
    double elapsed_ms() const { return timer_.elapsed() * 1e3; }
    unsigned int get_current_iteration() const { return current_iteration_.load(); }
    unsigned int get_max_iterations() const { return max_iterations_; }
    real_type get_progress() const { return progress_.load(); }
    real_type get_last_frametime() const { return last_frametime_.load(); }

  private:
    std::atomic<bool> stopped_{false};
    std::atomic<unsigned int> current_iteration_{0};
    unsigned int max_iterations_;
    std::atomic<double> progress_{0.0};
    Timer timer_;
    std::atomic<double> last_frametime_{std::numeric_limits<double>::max()};
};

}  // namespace core
}  // namespace lbm

#endif  // LBM_CONTROL_HPP
