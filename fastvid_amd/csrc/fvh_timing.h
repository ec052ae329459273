// fvh_timing.h -- optional per-kernel HIP-event timing (internal).
//
// When enabled through fvh_timing_enable, every kernel launch of the codec is
// bracketed by two hipEvents recorded on the launch stream; fvh_timing_read
// waits for them and reports launches and total milliseconds per kernel.
#pragma once

#include <hip/hip_runtime.h>

namespace fvh {

bool timing_enabled();
void timing_begin(const char* name, hipStream_t stream);
void timing_end(hipStream_t stream);

struct TimedLaunch {
  bool on;
  hipStream_t stream;
  TimedLaunch(const char* name, hipStream_t s) : on(timing_enabled()), stream(s) {
    if (on) timing_begin(name, s);
  }
  ~TimedLaunch() {
    if (on) timing_end(stream);
  }
};

}  // namespace fvh

#define FVH_TIMED(name, stream) ::fvh::TimedLaunch fvh_timed_scope(name, stream)
