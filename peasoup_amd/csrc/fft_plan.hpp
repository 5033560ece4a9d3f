// fft_plan.hpp - FFT decomposition descriptors shared by the plan builder,
// the batched search and the autotuner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "psx.h"

namespace psx {
namespace fft {

// A complex length M = L[0] * L[1] * ... with per-pass column tiles C[p].
// npass == 1: one contiguous pass.  npass == 2: pass 0 strided (length M2),
// pass 1 contiguous (length M1).  npass == 3: Q x M2 x M1.
struct Decomp {
  int npass = 0;
  int L[3] = {0, 0, 0};
  int C[3] = {0, 0, 0};
};

bool valid_complex_len(uint64_t m);
bool default_decomp(uint64_t M, Decomp& d);
bool pick_tiles(Decomp& d);
bool decomp_supported(const Decomp& d);
int plan_create_with(uint64_t n, int kind, uint32_t batch, const Decomp* forced, psx_fft_plan** out);
int plan_exec(psx_fft_plan* p, const void* in, void* out, hipStream_t s);

}  // namespace fft
}  // namespace psx
