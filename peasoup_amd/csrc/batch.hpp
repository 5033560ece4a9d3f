// batch.hpp - device entry points of the fused batched search.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "fft_plan.hpp"

namespace psx {
namespace batch {

// One above-threshold bin: trial slot, harmonic number (0 = fundamental),
// bin index and S/N.
struct Cand16 {
  uint32_t slot;
  uint32_t nh;
  int32_t idx;
  float snr;
};

// PeakFinder windows per harmonic number nh = 0..F (start inclusive, end
// exclusive) and the bin range [lo, hi) covering all of them.
struct Windows {
  int32_t start[6];
  int32_t end[6];
  int64_t lo, hi;
  float thr;
};

int launch_spec_post(const float2* Z, float* P, uint64_t M, int64_t pstride, const psx::fft::RootTable& rt,
                     const psx::fft::TrialDesc* d_descs, uint32_t ntr, hipStream_t s);
// What the row-layout spectrum pass also produces for the bound-filtered
// harmonic pass: Q (see launch_pool_peaks) and the fundamental's
// above-threshold bins, appended to out.
struct SpecBound {
  uint16_t* Q;
  int64_t qstride;
  Windows w;
  Cand16* out;
  uint32_t cap;
  uint32_t* count;
};
// Spectrum pass over the transposed layout of a row-writing last FFT pass
// (row r of L elements holds Z[r + R*e]); same output as launch_spec_post.
// bound (optional): writes Q and the fundamental's candidates instead of P
// (which may then be nullptr), as launch_pool_peaks would from P; its bound
// entries may be one bfloat16 step above those of the exact spectra.
int launch_spec_rows(const float2* Y, float* P, uint32_t R, uint32_t L, int64_t pstride,
                     const psx::fft::RootTable& rt, const psx::fft::TrialDesc* d_descs, uint32_t ntr, hipStream_t s,
                     const SpecBound* bound = nullptr);
// Bins per wave of the harmonic pass (8 per lane: measured 9% faster than
// 16 and 7% faster than 4 at 2^23).  It reads up to kHarmWaveBins - 1
// floats past the window end of a spectrum row, so a spectrum buffer of B
// rows of pstride floats is allocated with kHarmWaveBins floats of slack.
constexpr uint32_t kHarmWaveBins = 512;
int launch_harm_peaks(const float* P, int64_t pstride, const Windows& w, uint32_t nfolds,
                      const psx::fft::TrialDesc* d_descs, uint32_t ntr, Cand16* out, uint32_t cap, uint32_t* count,
                      hipStream_t s);
// Bound-filtered harmonic pass.  Q holds, per trial row of qstride 16-bit
// entries, bfloat16 upper bounds of bin pairs (entry u >= P[2u] and
// P[2u+1], rounded up); launch_pool_peaks writes it from P together with
// the fundamental's above-threshold bins, launch_harm_bound appends the
// folds' (nh >= 1) from Q, summing P exactly only where a bound crosses.
constexpr uint32_t kBoundPool = 2;         // bins per entry of Q
constexpr uint32_t kBoundWaveBins = 512;   // bins per wave of k_harm_bound
// The largest float v with fold_scale<k>(v) <= thr: fold k's running sum
// crosses the threshold iff it exceeds this value.
float fold_threshold_of(int k, float thr);
// Row stride (entries) of Q for spectrum rows of pstride floats.
int64_t bound_stride(int64_t pstride);
int launch_pool_peaks(const float* P, int64_t pstride, uint64_t nb, const Windows& w, uint16_t* Q, int64_t qstride,
                      const psx::fft::TrialDesc* d_descs, uint32_t ntr, Cand16* out, uint32_t cap, uint32_t* count,
                      hipStream_t s);
// The row-layout FFT output Y (R rows of L, trial stride R*L) from which
// launch_harm_bound recomputes the bins it sums exactly when the spectrum
// pass wrote no P.
struct RowsSource {
  const float2* Y;
  uint32_t R, L;
  psx::fft::RootTable rt;
};
// The exact sums read Y through a copy of rows in device memory (d_src,
// kRowsSourceBytes, written once by write_rows_source), which keeps those
// parameters out of the scalar registers of the kernel's main loop.
constexpr size_t kRowsSourceBytes = 64;
int write_rows_source(const RowsSource& rs, void* d_src);
int launch_harm_bound(const float* P, int64_t pstride, const uint16_t* Q, int64_t qstride, const Windows& w,
                      uint32_t nfolds, const psx::fft::TrialDesc* d_descs, uint32_t ntr, Cand16* out, uint32_t cap,
                      uint32_t* count, hipStream_t s, const RowsSource* rows = nullptr,
                      const void* d_rows_src = nullptr);
int launch_zero(uint32_t* p, hipStream_t s);

}  // namespace batch
}  // namespace psx
