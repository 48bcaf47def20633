/*
 * run_timer.hpp -- CPU clock ticks, CPU seconds and wall seconds of a timed region, the three
 * figures the miniapp drivers report per pass and for the whole run.
 */
#ifndef MINIISO_APPS_RUN_TIMER_HPP
#define MINIISO_APPS_RUN_TIMER_HPP

#include <chrono>
#include <ctime>

namespace miso_apps {

class RunTimer {
public:
    RunTimer() : c0_(std::clock()), w0_(std::chrono::steady_clock::now()) {}
    void stop()
    {
        c1_ = std::clock();
        w1_ = std::chrono::steady_clock::now();
    }
    long ticks() const { return (long)(c1_ - c0_); }
    double cpu_seconds() const { return (double)(c1_ - c0_) / CLOCKS_PER_SEC; }
    double wall_seconds() const { return std::chrono::duration<double>(w1_ - w0_).count(); }

private:
    std::clock_t c0_, c1_ = 0;
    std::chrono::steady_clock::time_point w0_, w1_;
};

}  // namespace miso_apps

#endif
