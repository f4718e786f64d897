// Optional per-kernel timing.  While timing is enabled, a TimedScope hands
// a start/stop event pair to the launch inside it (fmvd::launch), which
// attaches them to the kernel's dispatch packet (hipExtLaunchKernelGGL): the
// events then bracket the kernel's execution itself, without the gap a
// separate event marker would add before a short kernel.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <type_traits>

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
    // Never throws (it may run while an exception unwinds): a failure to
    // drain the recorded events is kept and reported by the next timing read.
    ~TimedScope() noexcept;
    TimedScope(const TimedScope&) = delete;
    TimedScope& operator=(const TimedScope&) = delete;

private:
    int kind_;
    hipStream_t s_;
    bool on_;
    hipEvent_t a_ = nullptr, b_ = nullptr;
};

static_assert(std::is_nothrow_destructible_v<TimedScope>);

// The event pair of the innermost open TimedScope on this thread, taken by
// the first launch inside it (both null when none is pending).
void take_launch_events(hipEvent_t& start, hipEvent_t& stop);

// Kernel launch that picks up a pending TimedScope's events.
template <typename... KArgs, typename... Args>
inline void launch(void (*kernel)(KArgs...), dim3 grid, dim3 block, unsigned shmem, hipStream_t s, Args... args)
{
    hipEvent_t a, b;
    take_launch_events(a, b);
    if (a) hipExtLaunchKernelGGL(kernel, grid, block, shmem, s, a, b, 0, static_cast<KArgs>(args)...);
    else kernel<<<grid, block, shmem, s>>>(static_cast<KArgs>(args)...);
}

}  // namespace fmvd
