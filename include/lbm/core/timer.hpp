// Wall-clock timer (reference: include/lbm/core/timer.hpp:24-47).
#ifndef LBM_TIMER_HPP
#define LBM_TIMER_HPP

#include <chrono>

namespace lbm
{
namespace core
{

class Timer
{
  public:
    Timer() : start_(clock::now()) {}
    void restart() { start_ = clock::now(); }
    // seconds since construction or the last restart()
    double elapsed() const { return std::chrono::duration<double>(clock::now() - start_).count(); }

  private:
    using clock = std::chrono::steady_clock;
    clock::time_point start_;
};

}  // namespace core
}  // namespace lbm

#endif  // LBM_TIMER_HPP
