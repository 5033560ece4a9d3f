// synth.hip - deterministic synthetic dedispersed time series on the device.
//
// Not part of the reference; it feeds tests and benchmarks with series of the
// configured shapes without host-side generation.  Sample i of trial d is
//   N(0,1) white noise          Philox4x32-10, key (seed, d), counter i/2
//   + red noise                 random walk with step sigma
//                               red_amp * 2*pi*0.1*tsamp (PSD = red_amp^2 x
//                               the white level at 0.1 Hz)
//   + pulsars                   Gaussian pulses (FWHM = duty * P) at phase
//                               (s - a/(2c) s (s - T))/P + phase0, the inverse
//                               of the resampler's warp, with per-sample
//                               amplitude snr / sqrt(n * sigma_phi * sqrt(pi))
// and the series mean is subtracted.
#include <cmath>

#include "common.hpp"

using psx::as_stream;

namespace {

constexpr unsigned kB = 256;
constexpr unsigned kPer = 16;
constexpr unsigned kTile = kB * kPer;
constexpr int kMaxPsr = 8;

struct Key {
  uint32_t k0, k1;
};
struct Ctr {
  uint32_t c0, c1, c2, c3;
};

// Philox4x32 (Salmon et al., SC'11), 10 rounds.
__device__ inline Ctr philox(Ctr c, Key k) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.c0), lo0 = M0 * c.c0;
    const uint32_t hi1 = __umulhi(M1, c.c2), lo1 = M1 * c.c2;
    c = Ctr{hi1 ^ c.c1 ^ k.k0, lo1, hi0 ^ c.c3 ^ k.k1, lo0};
    k.k0 += W0;
    k.k1 += W1;
  }
  return c;
}

__device__ inline double u53(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}

// Two N(0,1) draws for sample pair q of stream `which`.
__device__ inline double2 gauss2(uint64_t seed, uint32_t trial, uint32_t which, uint64_t q) {
  const Key k{(uint32_t)seed ^ (trial * 0x85EBCA6Bu), (uint32_t)(seed >> 32) ^ 0xC2B2AE35u};
  const Ctr r = philox(Ctr{(uint32_t)q, (uint32_t)(q >> 32), which, trial}, k);
  const double u1 = 1.0 - u53(r.c0, r.c1);  // (0, 1]
  const double u2 = u53(r.c2, r.c3);
  const double rad = sqrt(-2.0 * log(u1));
  double s, c;
  sincospi(2.0 * u2, &s, &c);
  return make_double2(rad * c, rad * s);
}

__device__ inline double gauss_at(uint64_t seed, uint32_t trial, uint32_t which, uint64_t i) {
  const double2 g = gauss2(seed, trial, which, i >> 1);
  return (i & 1) ? g.y : g.x;
}

struct PsrSet {
  int n;
  double inv_period[kMaxPsr];
  double kfac[kMaxPsr];  // a / (2c)
  double amp[kMaxPsr];
  double phase0[kMaxPsr];
  double inv_sigma[kMaxPsr];
};

__device__ inline double pulses(const PsrSet& ps, double s, double T) {
  double v = 0.0;
  for (int p = 0; p < ps.n; ++p) {
    const double tau = s - ps.kfac[p] * s * (s - T);
    const double ph = tau * ps.inv_period[p] + ps.phase0[p];
    const double d = (ph - rint(ph)) * ps.inv_sigma[p];
    if (fabs(d) < 8.0) v += ps.amp[p] * exp(-0.5 * d * d);
  }
  return v;
}

// Per-tile sums of the random-walk increments.
__global__ void __launch_bounds__(kB) k_walk_tilesum(uint64_t n, uint64_t seed, uint32_t trial, double sig,
                                                     double* __restrict__ tsum) {
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPer;
  double s = 0.0;
  for (unsigned q = 0; q < kPer; ++q)
    if (base + q < n) s += sig * gauss_at(seed, trial, 1u, base + q);
  __shared__ double red[kB];
  red[threadIdx.x] = s;
  __syncthreads();
  for (unsigned o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) tsum[blockIdx.x] = red[0];
}

// Exclusive scan of the tile sums in one block (sequential chunks per thread).
__global__ void __launch_bounds__(1024) k_walk_scan(double* __restrict__ tsum, unsigned ntiles) {
  __shared__ double part[1024];
  const unsigned per = (ntiles + 1023) / 1024;
  const unsigned lo = threadIdx.x * per;
  double s = 0.0;
  for (unsigned i = lo; i < lo + per && i < ntiles; ++i) s += tsum[i];
  part[threadIdx.x] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double acc = 0.0;
    for (unsigned t = 0; t < 1024; ++t) {
      const double v = part[t];
      part[t] = acc;
      acc += v;
    }
  }
  __syncthreads();
  double acc = part[threadIdx.x];
  for (unsigned i = lo; i < lo + per && i < ntiles; ++i) {
    const double v = tsum[i];
    tsum[i] = acc;
    acc += v;
  }
}

// White + walk + pulses; writes float samples and per-block double sums.
__global__ void __launch_bounds__(kB) k_synth(float* __restrict__ out, uint64_t n, double tsamp, uint64_t seed,
                                              uint32_t trial, double sig, const double* __restrict__ toff,
                                              PsrSet ps, double* __restrict__ bsum) {
  const uint64_t base = (uint64_t)blockIdx.x * kTile + (uint64_t)threadIdx.x * kPer;
  __shared__ double scan[kB];
  double walk[kPer];
  double run = 0.0;
  for (unsigned q = 0; q < kPer; ++q) {
    if (sig != 0.0 && base + q < n) run += sig * gauss_at(seed, trial, 1u, base + q);
    walk[q] = run;
  }
  scan[threadIdx.x] = run;
  __syncthreads();
  // inclusive Hillis-Steele scan of the thread totals
  for (unsigned o = 1; o < kB; o <<= 1) {
    const double v = threadIdx.x >= o ? scan[threadIdx.x - o] : 0.0;
    __syncthreads();
    scan[threadIdx.x] += v;
    __syncthreads();
  }
  const double before = (threadIdx.x ? scan[threadIdx.x - 1] : 0.0) + (sig != 0.0 ? toff[blockIdx.x] : 0.0);
  const double T = (double)n * tsamp;
  double local = 0.0;
  for (unsigned q = 0; q < kPer; ++q) {
    const uint64_t i = base + q;
    if (i >= n) break;
    double v = gauss_at(seed, trial, 0u, i) + before + walk[q];
    if (ps.n) v += pulses(ps, (double)i * tsamp, T);
    const float f = (float)v;
    out[i] = f;
    local += (double)f;
  }
  __syncthreads();
  scan[threadIdx.x] = local;
  __syncthreads();
  for (unsigned o = kB / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) scan[threadIdx.x] += scan[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) bsum[blockIdx.x] = scan[0];
}

__global__ void __launch_bounds__(1024) k_total(const double* __restrict__ bsum, unsigned nb, double* __restrict__ tot) {
  __shared__ double red[1024];
  double s = 0.0;
  for (unsigned i = threadIdx.x; i < nb; i += 1024) s += bsum[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (unsigned o = 512; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) tot[0] = red[0];
}

__global__ void __launch_bounds__(kB) k_sub_mean(float* __restrict__ out, uint64_t n, const double* __restrict__ tot) {
  const float m = (float)(tot[0] / (double)n);
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] -= m;
}

}  // namespace

extern "C" uint64_t psx_synth_work_bytes(uint64_t n) {
  const uint64_t ntiles = (n + kTile - 1) / kTile;
  return (2 * ntiles + 2) * sizeof(double);
}

extern "C" int psx_synth_series(float* out, uint64_t n, double tsamp, uint64_t seed, uint32_t trial, double red_amp,
                                 const psx_pulsar* pulsars, int npulsars, void* work, psx_stream s) {
  PSX_REQUIRE(out && work && n > 0 && tsamp > 0, "psx_synth_series: bad arguments");
  PSX_REQUIRE(npulsars >= 0 && npulsars <= kMaxPsr && (npulsars == 0 || pulsars),
              "psx_synth_series: 0..%d pulsars", kMaxPsr);
  const uint64_t ntiles64 = (n + kTile - 1) / kTile;
  PSX_REQUIRE(ntiles64 <= 0x7FFFFFFFull, "psx_synth_series: n too large");
  const unsigned ntiles = (unsigned)ntiles64;
  double* tsum = (double*)work;
  double* bsum = tsum + ntiles;
  double* tot = bsum + ntiles;
  PsrSet ps = {};
  ps.n = npulsars;
  for (int p = 0; p < npulsars; ++p) {
    const psx_pulsar& q = pulsars[p];
    PSX_REQUIRE(q.period > 0 && q.duty > 0 && q.duty < 1, "psx_synth_series: bad pulsar %d", p);
    const double sigma = q.duty / (2.0 * sqrt(2.0 * log(2.0)));
    ps.inv_period[p] = 1.0 / q.period;
    ps.kfac[p] = q.accel / (2.0 * 299792458.0);
    ps.amp[p] = q.snr / sqrt((double)n * sigma * sqrt(M_PI));
    ps.phase0[p] = q.phase0;
    ps.inv_sigma[p] = 1.0 / sigma;
  }
  const double sig = red_amp > 0 ? red_amp * 2.0 * M_PI * 0.1 * tsamp : 0.0;
  const hipStream_t hs = as_stream(s);
  if (sig != 0.0) {
    k_walk_tilesum<<<ntiles, kB, 0, hs>>>(n, seed, trial, sig, tsum);
    PSX_CHECK_LAUNCH("k_walk_tilesum");
    k_walk_scan<<<1, 1024, 0, hs>>>(tsum, ntiles);
    PSX_CHECK_LAUNCH("k_walk_scan");
  }
  k_synth<<<ntiles, kB, 0, hs>>>(out, n, tsamp, seed, trial, sig, tsum, ps, bsum);
  PSX_CHECK_LAUNCH("k_synth");
  k_total<<<1, 1024, 0, hs>>>(bsum, ntiles, tot);
  PSX_CHECK_LAUNCH("k_total");
  k_sub_mean<<<psx::grid_for(n, kB), kB, 0, hs>>>(out, n, tot);
  PSX_CHECK_LAUNCH("k_sub_mean");
  return psx::ok();
}
