// Run-level accounting shared by the SIMPLE and PISO drivers: the
// iteration-compute window of RunTime::recordIteration
// (src/runtime/runtime.cpp:117-130) and the host reads of device reductions.
#pragma once

#include <chrono>
#include <cstdint>

namespace bsim {

struct RunClock {
    using Clock = std::chrono::steady_clock;
    Clock::time_point created = Clock::now(), firstEnd;
    bool started = false;
    std::uint64_t intervals = 0;  // reported_iteration_intervals
    double window = 0.0;          // iteration_compute_seconds: end of the first iteration to the last
    std::uint64_t reductions = 0; // device reductions read by the host (convergence measures, residuals)
    // Called at the end of every outer iteration or time step.
    void record() {
        const auto t = Clock::now();
        if (!started) {
            firstEnd = t;
            started = true;
            return;
        }
        ++intervals;
        window = std::chrono::duration<double>(t - firstEnd).count();
    }
    double elapsed() const { return std::chrono::duration<double>(Clock::now() - created).count(); }
};

}  // namespace bsim
