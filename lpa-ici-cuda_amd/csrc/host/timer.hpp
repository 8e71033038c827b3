// Wall-clock timer started at construction (src/Timer.hpp: std::chrono::steady_clock,
// seconds as double).
#ifndef LPA_ICI_TIMER_HPP
#define LPA_ICI_TIMER_HPP

#include <chrono>

class Timer {
  public:
    void reset() { start_ = std::chrono::steady_clock::now(); }
    double elapsed() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
    }

  private:
    std::chrono::steady_clock::time_point start_ = std::chrono::steady_clock::now();
};

#endif
