// reduce_async.hpp - device-resident forms of the prologue's
// host-synchronising reductions (psx_remove_baseline, psx_mean_rms), for a
// prologue that runs ahead on its own stream without host round trips.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psx {

// remove_baseline_async runs `steps` iterations of psx_remove_baseline's loop
// on the device (the same mean, subtraction and stopping rule; iterations
// after the stop do nothing); st->done == 0 afterwards means the loop had not
// stopped and the caller must finish it with psx_remove_baseline.
struct BaselineState {
  int32_t count;
  int32_t done;
};
int remove_baseline_async(float* x, uint64_t n, void* work, BaselineState* st, int steps, hipStream_t s);
// Sums (x, x^2) over [first, n) into *out (device), as psx_mean_rms reduces;
// nothing is summed (and *out is left as it is) when skip points to a
// stopped baseline loop.
int sum_sq_async(const float* x, uint64_t first, uint64_t n, void* work, double2* out, hipStream_t s,
                 const BaselineState* skip = nullptr);

}  // namespace psx
