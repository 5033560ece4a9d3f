// kernels_dev.hpp - device-side arithmetic shared by the plain per-stage
// kernels and the fused batched kernels.  Each helper reproduces one
// expression of the reference kernels (src/kernels.cu) with the same
// operation order and precision; multiply-adds that nvcc contracts in the
// reference are written as explicit fma, everything else is unfused (the
// library is built with -ffp-contract=off).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psx {

// device_resampleII (src/kernels.cu:361): (a*tsamp) is a float product,
// widened to double before the divide by 2c.
__host__ __device__ inline double accel_fact(float acc, float tsamp) {
  const float at = acc * tsamp;
  return (double)at / (2 * 299792458.0);
}

// getAcceleratedIndexII (src/kernels.cu:321-324): __double2ull_rn of
// id + id*af*(id - size), contracted to fma(id*af, id - size, id).  For
// |af|*n < 1 the value lies in [0, n]; the clamp only guards the n edge.
__device__ inline uint64_t resample_index(uint64_t i, uint64_t n, double af) {
  const double di = (double)i;
  double v = fma(di * af, di - (double)n, di);
  v = v > 0.0 ? v : 0.0;
  const uint64_t j = (uint64_t)rint(v);
  return j < n ? j : n - 1;
}

// The same index for series shorter than 2^31 samples, in 32-bit integers
// (identical values: every operand is exactly representable either way).
// n >= 1; the n edge clamps with one unsigned min.
__device__ inline uint32_t resample_index32(uint32_t i, uint32_t n, double dn, double af) {
  const double di = (double)i;
  double v = fma(di * af, di - dn, di);
  v = v > 0.0 ? v : 0.0;
  return min((uint32_t)rint(v), n - 1);
}

// bin_interbin_series_kernel (src/kernels.cu:250-257): the 0.5 is a double
// literal; halving in float gives the identical correctly rounded result.
__device__ inline float interbin(float2 x, float2 prev) {
  const float ampsq = fmaf(x.x, x.x, x.y * x.y);
  const float dr = x.x - prev.x;
  const float di = x.y - prev.y;
  const float ampsq_diff = 0.5f * fmaf(dr, dr, di * di);
  return sqrtf(fmaxf(ampsq, ampsq_diff));
}

// Order statistics for median_scrunch5 (src/kernels.cu:902-959).  The median
// is a unique value, so any correct network returns the reference's value.
__device__ inline void cswap(float& a, float& b) {
  const float lo = fminf(a, b);
  const float hi = fmaxf(a, b);
  a = lo;
  b = hi;
}
__device__ inline float median3(float a, float b, float c) {
  cswap(a, b);
  cswap(b, c);
  cswap(a, b);
  return b;
}
__device__ inline float median4(float a, float b, float c, float d) {
  cswap(a, b);
  cswap(c, d);
  cswap(a, c);
  cswap(b, d);
  cswap(b, c);
  return 0.5f * (b + c);
}
// Seven compare-exchanges leave the median of five in c.
__device__ inline float median5(float a, float b, float c, float d, float e) {
  cswap(a, b);
  cswap(d, e);
  cswap(a, d);
  cswap(b, e);
  cswap(b, c);
  cswap(c, d);
  cswap(b, c);
  return c;
}

// linear_stretch_functor::operator() (src/kernels.cu:1018-1023).  x is
// float(i)*step; the neighbour read past the last input (undefined in the
// reference once x rounds above in_count-1) is clamped to the last input.
__device__ inline float stretch_at(const float* __restrict__ in, uint64_t in_count, unsigned i, float step) {
  const float x = (float)i * step;
  unsigned j = (unsigned)x;
  const float fr = x - (float)j;
  const uint64_t last = in_count - 1;
  const float a = in[j < last ? j : last];
  const float b = in[j + 1 < last ? j + 1 : last];
  return (fr > 1e-5f) ? fmaf(fr, b - a, a) : a + 0.f;
}

// Scale of fold K's running sum (1-based): times 2^-K/2, computed in double
// for odd K as the reference does (src/kernels.cu:38-98).  Host and device
// round every step alike (IEEE multiply and conversion, nearest even).
template <int K>
__host__ __device__ inline float fold_scale(float val) {
  if constexpr (K == 1) return (float)((double)val * 0.70710678118654752440);
  else if constexpr (K == 2) return val * 0.5f;
  else if constexpr (K == 3) return (float)((double)val * 0.35355339059327376220);
  else if constexpr (K == 4) return val * 0.25f;
  else return (float)((double)val * 0.17677669529663688110);
}

// harmonic_sum_kernel (src/kernels.cu:38-98).  Fold k adds the bins
// (int)(i*j/2^(k+1) + 0.5) for odd j, in the order the reference lists them
// (fold 1 adds 3/4 before 1/4), to one running float sum, and stores the
// running sum times 2^-(k+1)/2 computed in double.
template <int NF>
__device__ inline void harmonic_sums(const float* __restrict__ P, uint64_t i, float (&s)[5]) {
  float val = P[i];
  if (NF > 0) {
    val += P[(i + 1) >> 1];
    s[0] = fold_scale<1>(val);
  }
  if (NF > 1) {
    val += P[(3 * i + 2) >> 2];
    val += P[(i + 2) >> 2];
    s[1] = fold_scale<2>(val);
  }
  if (NF > 2) {
#pragma unroll
    for (uint64_t j = 1; j < 8; j += 2) val += P[(j * i + 4) >> 3];
    s[2] = fold_scale<3>(val);
  }
  if (NF > 3) {
#pragma unroll
    for (uint64_t j = 1; j < 16; j += 2) val += P[(j * i + 8) >> 4];
    s[3] = fold_scale<4>(val);
  }
  if (NF > 4) {
#pragma unroll
    for (uint64_t j = 1; j < 32; j += 2) val += P[(j * i + 16) >> 5];
    s[4] = fold_scale<5>(val);
  }
}

// The same sums with 32-bit bin arithmetic (valid for i < 2^27, where every
// j*i + 2^(m-1) fits in 32 bits).
template <int NF>
__device__ inline void harmonic_sums32(const float* __restrict__ P, uint32_t i, float (&s)[5]) {
  float val = P[i];
  if (NF > 0) {
    val += P[(i + 1) >> 1];
    s[0] = fold_scale<1>(val);
  }
  if (NF > 1) {
    val += P[(3 * i + 2) >> 2];
    val += P[(i + 2) >> 2];
    s[1] = fold_scale<2>(val);
  }
  if (NF > 2) {
#pragma unroll
    for (uint32_t j = 1; j < 8; j += 2) val += P[(j * i + 4) >> 3];
    s[2] = fold_scale<3>(val);
  }
  if (NF > 3) {
#pragma unroll
    for (uint32_t j = 1; j < 16; j += 2) val += P[(j * i + 8) >> 4];
    s[3] = fold_scale<4>(val);
  }
  if (NF > 4) {
#pragma unroll
    for (uint32_t j = 1; j < 32; j += 2) val += P[(j * i + 16) >> 5];
    s[4] = fold_scale<5>(val);
  }
}

// Number of set bits of mask in the lanes below the calling lane (wave64).
__device__ inline unsigned mbcnt64(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

}  // namespace psx
