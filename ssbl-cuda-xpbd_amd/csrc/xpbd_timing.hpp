// xpbd_timing.hpp -- optional device-side timing of every kernel launch.
//
// When enabled (ssbl_ext_kernel_timing), every launch made through launch_k
// inside a TimedLaunch scope carries a start and a stop event that the
// dispatch itself stamps (hipExtLaunchKernel), so a duration is the kernel's
// own execution time, without the queue gaps around it.  Durations are
// resolved and summed per scope name by ssbl_ext_kernel_stats.  Disabled, a
// launch costs one relaxed atomic load more than a plain launch.
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <cstdint>

namespace ssbl {

bool timing_enabled();
void timing_take(hipEvent_t* begin, hipEvent_t* end);
void timing_push(const char* name, hipEvent_t begin, hipEvent_t end);
const char*& timing_scope_slot();

// Names the launches made while it is alive.
class TimedLaunch {
public:
    TimedLaunch(const char* name, hipStream_t) : saved_(timing_scope_slot()) { timing_scope_slot() = name; }
    ~TimedLaunch() { timing_scope_slot() = saved_; }
    TimedLaunch(const TimedLaunch&) = delete;
    TimedLaunch& operator=(const TimedLaunch&) = delete;

private:
    const char* saved_;
};

template <typename F, typename... Args>
inline void launch_k(F kernel, const dim3& grid, const dim3& block, std::uint32_t shmem, hipStream_t stream, Args... args) {
    const char* name = timing_enabled() ? timing_scope_slot() : nullptr;
    hipEvent_t begin = nullptr, end = nullptr;
    if (name) timing_take(&begin, &end);
    hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, begin, end, 0u, args...);
    if (name) timing_push(name, begin, end);
}

}  // namespace ssbl
