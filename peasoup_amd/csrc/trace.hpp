// trace.hpp - named host ranges for the profiler's marker trace (roctx).
// The ranges carry the reference's NVTX range names ("DM-Loop",
// "Acceleration-Loop", "Harmonic summing", "Dedisperse":
// src/pipeline_multi.cu:154,244,473, include/transforms/harmonicfolder.hpp:28)
// so timelines of both builds read alike.  Without a profiler attached a
// range costs two calls into librocprofiler-sdk-roctx that return at once.
#pragma once
#include <rocprofiler-sdk-roctx/roctx.h>

namespace psx {

class TraceRange {
 public:
  explicit TraceRange(const char* name) { roctxRangePushA(name); }
  ~TraceRange() { roctxRangePop(); }
  TraceRange(const TraceRange&) = delete;
  TraceRange& operator=(const TraceRange&) = delete;
};

}  // namespace psx
