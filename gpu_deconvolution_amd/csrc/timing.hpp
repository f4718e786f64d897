// Optional per-kernel timing: HIP events recorded on the launch stream
// around each kernel launch while timing is enabled.
#pragma once
#include <hip/hip_runtime.h>

namespace fmvd {

enum TimingKind {
    kTimeColumn = 0,
    kTimeRatio = 1,
    kTimeUpdate = 2,
    kTimeIngest = 3,
    kTimeEgress = 4,
    kTimeTransform = 5,
    kTimingKinds = 6
};

bool timing_on();

// sample = false skips this launch: callers that issue many short launches
// time a regular subset of them, since every recorded event pair costs
// microseconds of GPU time between kernels.
class TimedScope {
public:
    TimedScope(int kind, hipStream_t s, bool sample = true);
    ~TimedScope() noexcept(false);
    TimedScope(const TimedScope&) = delete;
    TimedScope& operator=(const TimedScope&) = delete;

private:
    int kind_;
    hipStream_t s_;
    bool on_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

}  // namespace fmvd
