// batch.hip - fused kernels of the batched acceleration search (gfx950).
//
// One batch = up to B acceleration trials of the same FFT length.  Per batch:
//   k_fft_pass (resample fused into the first pass) ... FFT passes -> Z
//   k_spec_post   untangle + interbin + normalise          Z -> P
//   k_harm_peaks  2..2^F harmonic sums + threshold +
//                 wave-ballot compaction                   P -> candidates
// The per-trial arithmetic is the plain path's, operation for operation
// (untangled_bin, interbin, (v-mean)/std, harmonic_sums), so every value the
// batched path thresholds equals the plain path's bit for bit; only the
// intermediate HBM round trips (X, the interbinned spectrum before
// normalisation, the F fold spectra) disappear.
#include "batch.hpp"
#include "common.hpp"
#include "kernels_dev.hpp"

using namespace psx::fft;

namespace psx {
namespace batch {
namespace {

constexpr unsigned kB = 256;

// untangle + form_interpolated + normalise (src/kernels.cu:238-259,496-507)
// for bins 2q and 2q+1 of trial blockIdx.y.
__global__ void __launch_bounds__(kB) k_spec_post(const float2* __restrict__ Z, float* __restrict__ P, uint64_t M,
                                                  int64_t pstride, RootTable rt, const TrialDesc* __restrict__ descs) {
  const uint32_t tr = blockIdx.y;
  const float2* z = Z + (int64_t)tr * (int64_t)M;
  float* p = P + (int64_t)tr * pstride;
  const float mean_n = descs[tr].mean_n;
  const float std_n = descs[tr].std_n;
  const uint64_t nb = M + 1;
  const uint64_t npair = (nb + 1) / 2;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < npair; q += stride) {
    const uint64_t k = 2 * q;
    const float2 xm1 = k > 0 ? untangled_bin(z, M, k - 1, rt) : make_float2(0.f, 0.f);
    const float2 x0 = untangled_bin(z, M, k, rt);
    float v0 = psx::interbin(x0, xm1);
    v0 -= mean_n;
    v0 /= std_n;
    if (k + 1 < nb) {
      const float2 x1 = untangled_bin(z, M, k + 1, rt);
      float v1 = psx::interbin(x1, x0);
      v1 -= mean_n;
      v1 /= std_n;
      reinterpret_cast<float2*>(p)[q] = make_float2(v0, v1);
    } else {
      p[k] = v0;
    }
  }
}

// harmonic_sum_kernel (src/kernels.cu:34-100) + device_find_peaks
// (src/kernels.cu:391-423) for the fundamental and every fold, each over
// its own PeakFinder window (peakfinder.hpp:78-83).  Survivors are appended
// to `out` with one atomic per wave; the host orders them by bin.
template <int NF>
__global__ void __launch_bounds__(kB) k_harm_peaks(const float* __restrict__ P, int64_t pstride, Windows w,
                                                   const TrialDesc* __restrict__ descs, Cand16* __restrict__ out,
                                                   uint32_t cap, uint32_t* __restrict__ count) {
  const uint32_t tr = blockIdx.y;
  const float* p = P + (int64_t)tr * pstride;
  const uint32_t slot = descs[tr].slot;
  const unsigned lane = threadIdx.x & 63u;
  const int64_t step = (int64_t)gridDim.x * kB;
  for (int64_t base = w.lo + (int64_t)blockIdx.x * kB; base < w.hi; base += step) {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < w.hi;
    float s[6];
    s[0] = 0.f;
    if (valid) {
      float f[5];
      psx::harmonic_sums<NF>(p, (uint64_t)i, f);
      s[0] = p[i];
#pragma unroll
      for (int k = 0; k < NF; ++k) s[k + 1] = f[k];
    }
#pragma unroll
    for (int nh = 0; nh <= NF; ++nh) {
      const bool hit = valid && i >= w.start[nh] && i < w.end[nh] && s[nh] > w.thr;
      const uint64_t mask = __ballot(hit);
      if (mask) {
        const int leader = __ffsll((long long)mask) - 1;
        uint32_t b = 0;
        if ((int)lane == leader) b = atomicAdd(count, (uint32_t)__popcll(mask));
        b = __shfl(b, leader);
        if (hit) {
          const uint32_t pos = b + psx::mbcnt64(mask);
          if (pos < cap) out[pos] = Cand16{slot, (uint32_t)nh, (int32_t)i, s[nh]};
        }
      }
    }
  }
}

__global__ void k_zero_u32(uint32_t* p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0u;
}

}  // namespace

int launch_spec_post(const float2* Z, float* P, uint64_t M, int64_t pstride, const RootTable& rt,
                     const TrialDesc* d_descs, uint32_t ntr, hipStream_t s) {
  const uint64_t npair = (M + 2) / 2;
  const unsigned gx = psx::grid_for(npair, kB, 2048);
  k_spec_post<<<dim3(gx, ntr), kB, 0, s>>>(Z, P, M, pstride, rt, d_descs);
  PSX_CHECK_LAUNCH("k_spec_post");
  return PSX_OK;
}

int launch_harm_peaks(const float* P, int64_t pstride, const Windows& w, uint32_t nfolds, const TrialDesc* d_descs,
                      uint32_t ntr, Cand16* out, uint32_t cap, uint32_t* count, hipStream_t s) {
  if (w.hi <= w.lo) return PSX_OK;
  const unsigned gx = psx::grid_for((uint64_t)(w.hi - w.lo), kB, 2048);
  const dim3 g(gx, ntr);
  switch (nfolds) {
    case 1: k_harm_peaks<1><<<g, kB, 0, s>>>(P, pstride, w, d_descs, out, cap, count); break;
    case 2: k_harm_peaks<2><<<g, kB, 0, s>>>(P, pstride, w, d_descs, out, cap, count); break;
    case 3: k_harm_peaks<3><<<g, kB, 0, s>>>(P, pstride, w, d_descs, out, cap, count); break;
    case 4: k_harm_peaks<4><<<g, kB, 0, s>>>(P, pstride, w, d_descs, out, cap, count); break;
    case 5: k_harm_peaks<5><<<g, kB, 0, s>>>(P, pstride, w, d_descs, out, cap, count); break;
    default: return psx::fail(PSX_ERR_ARG, "nfolds must be 1..5");
  }
  PSX_CHECK_LAUNCH("k_harm_peaks");
  return PSX_OK;
}

int launch_zero(uint32_t* p, hipStream_t s) {
  k_zero_u32<<<1, 64, 0, s>>>(p);
  PSX_CHECK_LAUNCH("k_zero_u32");
  return PSX_OK;
}

}  // namespace batch
}  // namespace psx
