// xpbd_timing.hpp -- optional device-side timing of every kernel launch.
//
// When enabled (ssbl_ext_kernel_timing), a TimedLaunch brackets the launch
// made in its scope with two events on the launch's stream; durations are
// resolved and summed per kernel name by ssbl_ext_kernel_stats.  Disabled, a
// TimedLaunch costs one relaxed atomic load.
#pragma once

#include <hip/hip_runtime.h>

namespace ssbl {

bool timing_enabled();
void timing_record(const char* name, hipStream_t stream, hipEvent_t* begin_out);
void timing_finish(const char* name, hipStream_t stream, hipEvent_t begin);

class TimedLaunch {
public:
    TimedLaunch(const char* name, hipStream_t stream) : name_(name), stream_(stream) {
        if (timing_enabled()) { active_ = true; timing_record(name_, stream_, &begin_); }
    }
    ~TimedLaunch() {
        if (active_) timing_finish(name_, stream_, begin_);
    }
    TimedLaunch(const TimedLaunch&) = delete;
    TimedLaunch& operator=(const TimedLaunch&) = delete;

private:
    const char* name_;
    hipStream_t stream_;
    hipEvent_t begin_ = nullptr;
    bool active_ = false;
};

}  // namespace ssbl
