// fft_plan.hpp - FFT decomposition descriptors shared by the plan builder,
// the batched search and the autotuner.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_dev.hpp"
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
// Roots W_n^k of the real length n as a two-level table: W = hi[k>>shift]*lo[k&mask].
struct RootTable {
  const float2* lo;
  const float2* hi;
  uint32_t shift, mask;
};

__device__ inline float2 root(const RootTable& t, uint64_t k) { return cmul(t.hi[k >> t.shift], t.lo[k & t.mask]); }
// The same root for k < 2^29, loaded with 32-bit offsets from the tables.
__device__ inline float2 root32(const RootTable& t, uint32_t k) {
  return cmul(ldg2(t.hi, k >> t.shift), ldg2(t.lo, k & t.mask));
}

// R2C untangle of one bin: X = (zk + conj(zm))/2 - (i/2) w (zk - conj(zm)).
__device__ inline float2 untangle_one(float2 zk, float2 zm, float2 w) {
  const float2 e = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
  const float2 d = make_float2(zk.x - zm.x, zk.y + zm.y);
  const float2 wd = cmul(w, d);
  return make_float2(e.x + 0.5f * wd.y, e.y - 0.5f * wd.x);
}

// Bin k in [0, M] of the untangled spectrum, computed exactly as the in-place
// untangle pass does (so fused consumers see identical values).
__device__ inline float2 untangled_bin(const float2* __restrict__ z, uint64_t M, uint64_t k, const RootTable& rt) {
  if (k == 0) return untangle_one(z[0], z[0], root(rt, 0));
  if (k == M) return untangle_one(z[0], z[0], make_float2(-1.f, 0.f));
  if (2 * k <= M) return untangle_one(z[k], z[M - k], root(rt, k));
  const float2 w = root(rt, M - k);
  return untangle_one(z[k], z[M - k], make_float2(-w.x, w.y));
}

int plan_create_with(uint64_t n, int kind, uint32_t batch, const Decomp* forced, psx_fft_plan** out);
int plan_exec(psx_fft_plan* p, const void* in, void* out, hipStream_t s);
// Forward complex passes of an R2C plan whose first pass reads trial
// blockIdx.z through the resampler (descs); writes the half-length complex
// spectra Z (trial stride M) to zout without untangling.
// rows_out: the last pass writes the transposed layout instead (see
// rows_layout); only for plans that have it.
int exec_fused_r2c(psx_fft_plan* p, const TrialDesc* d_descs, uint32_t ntrials, float2* zout, hipStream_t s,
                   bool rows_out = false);
int exec_fused_pass(psx_fft_plan* p, int q, const TrialDesc* d_descs, uint32_t ntrials, float2* zout, hipStream_t s,
                    bool rows_out = false);
// True when a plan's last pass can write contiguous rows: row r of row_len
// elements (the last pass's length) then holds Z[r + rows*e],
// e = 0..row_len-1 (rows = the product of the other passes' lengths),
// instead of the natural Z[k].
bool rows_layout(const psx_fft_plan* p, uint32_t* rows, uint32_t* row_len);
int num_passes(const psx_fft_plan* p);
// True when pass q transforms the work buffer in place (the middle pass of
// three-pass plans): launching it again transforms its output again.
bool pass_in_place(const psx_fft_plan* p, int q);
RootTable root_table(const psx_fft_plan* p);
uint64_t complex_len(const psx_fft_plan* p);
const Decomp& plan_decomp(const psx_fft_plan* p);

}  // namespace fft
}  // namespace psx
