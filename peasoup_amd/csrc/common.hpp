// common.hpp - shared helpers for the libpsx HIP sources (gfx950).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>

#include "psx.h"

namespace psx {

// Records a thread-local error message and returns code.
int fail(int code, const char* fmt, ...);
// Clears the thread-local error message and returns PSX_OK.
int ok();

inline hipStream_t as_stream(psx_stream s) { return reinterpret_cast<hipStream_t>(s); }

// Grid size for a grid-stride loop over n items with `per_block` items per
// block-iteration; capped so one launch keeps ~16 blocks per CU in flight.
inline unsigned grid_for(uint64_t n, uint64_t per_block, unsigned cap = 4096) {
  uint64_t g = (n + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return (unsigned)g;
}

// Workgroups are dispatched round-robin over the 8 XCDs of an MI355X.  This
// maps the hardware block id to a logical id such that each XCD receives a
// contiguous run of logical ids, so neighbouring tiles share an L2.
__device__ inline unsigned xcd_logical_block(unsigned hw, unsigned nblocks) {
  const unsigned nx = 8;
  const unsigned full = (nblocks / nx) * nx;
  if (hw >= full) return hw;
  const unsigned per = full / nx;
  return (hw % nx) * per + hw / nx;
}

}  // namespace psx

#define PSX_CHECK_HIP(expr)                                                                   \
  do {                                                                                        \
    hipError_t psx_e_ = (expr);                                                               \
    if (psx_e_ != hipSuccess)                                                                 \
      return ::psx::fail(PSX_ERR_HIP, "%s failed: %s", #expr, hipGetErrorString(psx_e_));     \
  } while (0)

#define PSX_CHECK_LAUNCH(name)                                                                \
  do {                                                                                        \
    hipError_t psx_e_ = hipGetLastError();                                                    \
    if (psx_e_ != hipSuccess)                                                                 \
      return ::psx::fail(PSX_ERR_HIP, "%s launch failed: %s", name, hipGetErrorString(psx_e_)); \
  } while (0)

#define PSX_REQUIRE(cond, ...)                                                                \
  do {                                                                                        \
    if (!(cond)) return ::psx::fail(PSX_ERR_ARG, __VA_ARGS__);                                \
  } while (0)
