// batch.hip - fused kernels of the batched acceleration search (gfx950).
//
// One batch = up to B acceleration trials of the same FFT length.  Per batch:
//   k_fft_pass (resample fused into the first pass) ... FFT passes -> Z
//   k_spec_post   untangle + interbin + normalise          Z -> P
//   k_pool_peaks  fundamental threshold + bound spectrum   P -> Q, candidates
//   k_harm_bound  2..2^F harmonic sums over Q (an upper bound of every
//                 fold sum); exact sums of P where a bound
//                 crosses the threshold                    Q, P -> candidates
// (k_harm_peaks, the exact sums of every bin, is the alternative harmonic
// pass: same candidates, more memory traffic.)
// The per-trial arithmetic is the plain path's, operation for operation
// (untangled_bin, interbin, (v-mean)/std, harmonic_sums), so every value the
// batched path thresholds equals the plain path's bit for bit; only the
// intermediate HBM round trips (X, the interbinned spectrum before
// normalisation, the F fold spectra) disappear.
#include "batch.hpp"

#include <cmath>
#include <cstring>

#include "common.hpp"
#include "kernels_dev.hpp"

using namespace psx::fft;

namespace psx {
namespace batch {
namespace {

constexpr unsigned kB = 256;

// untangle + form_interpolated + normalise (src/kernels.cu:238-259,496-507).
// The half-length spectrum pair (Z[k], Z[M-k]), k <= M/2, untangles into both
// X[k] and X[M-k] (the two halves of the plain untangle pass, same
// arithmetic).  Interbinning needs X[k-1] beside X[k], and X[M-k-1] (the
// upper half of pair k+1) beside X[M-k].  A wave holds 64 consecutive pairs
// k = kw-1 .. kw+62 (one lane per pair, both loads coalesced, the mirror row
// descending) and takes every neighbour from the adjacent lane by a shuffle:
// lanes 1..63 write P[k], lanes 0..62 write P[M-k], and consecutive waves
// advance by 63 pairs, so no lane ever loads a neighbour itself.
__device__ inline float normalised(float2 x, float2 prev, float mean_n, float std_n) {
  float v = psx::interbin(x, prev);
  v -= mean_n;
  v /= std_n;
  return v;
}

// An upper bound of normalised(x, prev, mean_n, std_n) (std_n > 0) from the
// hardware's approximate square root and reciprocal (each within 1 ulp):
// the interbin square root is raised by 2^-20 of itself (8 ulp, more than
// the approximation's error), the mean subtracted as the reference does
// (rounding is monotone), and the quotient taken as a product with the
// approximate reciprocal and raised by 2^-20 of its magnitude (its error is
// under 2 ulp of the quotient, plus the half ulp of correct rounding).
// About a third of normalised's instructions: the bound-form spectrum pass
// needs the exact value only where this one exceeds the threshold.
__device__ inline float normalised_up(float2 x, float2 prev, float mean_n, float std_n) {
  const float ampsq = fmaf(x.x, x.x, x.y * x.y);
  const float dr = x.x - prev.x;
  const float di = x.y - prev.y;
  const float ampsq_diff = 0.5f * fmaf(dr, dr, di * di);
  const float a = __builtin_amdgcn_sqrtf(fmaxf(ampsq, ampsq_diff));
  const float v = fmaf(a, 0x1p-20f, a) - mean_n;
  const float q = v * __builtin_amdgcn_rcpf(std_n);
  return fmaf(fabsf(q), 0x1p-20f, q);
}

__device__ inline float2 shfl_f2(float2 v, int src) {
  return make_float2(__shfl(v.x, src), __shfl(v.y, src));
}

constexpr uint32_t kStep = 63;  // pairs written per wave

// bfloat16 bits (the top half of the float) of the smallest bfloat16 >= x:
// a positive float with any of its low 16 bits set rounds its top half up
// (a carry into the exponent is the next binade, and past the largest
// finite value +inf); a negative one truncates toward zero, which is up.
__device__ inline uint16_t bf16_up(float x) {
  const uint32_t u = __float_as_uint(x);
  return (uint16_t)(((int32_t)u < 0 ? u : u + 0xFFFFu) >> 16);
}
__device__ inline float bf16_value(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

// Reads the kBoundD LDS entries e[0], e[STRIDE], ... into the high halves of
// t[0..kBoundD), whose low halves the caller keeps at zero: ds_read_u16_d16_hi
// writes only the high half, so t[d] is the entry's float value with no
// conversion instruction.  The reads and their wait are one asm statement.
template <int STRIDE>
__device__ inline void lds_bf16x8(const uint16_t* e, float (&t)[8]) {
  const uint32_t a = (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) uint16_t*)e;
  asm volatile(
      "ds_read_u16_d16_hi %0, %8\n\t"
      "ds_read_u16_d16_hi %1, %8 offset:%9\n\t"
      "ds_read_u16_d16_hi %2, %8 offset:%10\n\t"
      "ds_read_u16_d16_hi %3, %8 offset:%11\n\t"
      "ds_read_u16_d16_hi %4, %8 offset:%12\n\t"
      "ds_read_u16_d16_hi %5, %8 offset:%13\n\t"
      "ds_read_u16_d16_hi %6, %8 offset:%14\n\t"
      "ds_read_u16_d16_hi %7, %8 offset:%15\n\t"
      "s_waitcnt lgkmcnt(0)"
      : "+v"(t[0]), "+v"(t[1]), "+v"(t[2]), "+v"(t[3]), "+v"(t[4]), "+v"(t[5]), "+v"(t[6]), "+v"(t[7])
      : "v"(a), "i"(2 * STRIDE), "i"(4 * STRIDE), "i"(6 * STRIDE), "i"(8 * STRIDE), "i"(10 * STRIDE),
        "i"(12 * STRIDE), "i"(14 * STRIDE));
}

// Outputs of the spectrum pass for the bound-filtered harmonic pass: the
// bound spectrum (one bfloat16 upper bound per bin pair 2u, 2u+1) and
// the fundamental's above-threshold bins (device_find_peaks over
// PeakFinder's nh = 0 window, src/kernels.cu:391-423, peakfinder.hpp:78-83).
struct BoundOut {
  uint16_t* Q;
  int64_t qstride;
  int32_t start0, end0;
  float thr;
  Cand16* out;
  uint32_t cap;
  uint32_t* count;
};

// Appends the calling lanes' hits (bin k, value v) as fundamental
// candidates of slot; every lane of the wave calls it, lane 0 included.
__device__ inline void append_fundamental(bool hit, uint32_t slot, uint32_t k, float v, const BoundOut& bo) {
  const uint64_t mask = __ballot(hit);
  if (mask == 0) return;
  uint32_t b = 0;
  if ((threadIdx.x & 63u) == 0) b = atomicAdd(bo.count, (uint32_t)__popcll(mask));
  b = __shfl(b, 0);
  if (hit) {
    const uint32_t pos = b + psx::mbcnt64(mask);
    if (pos < bo.cap) bo.out[pos] = Cand16{slot, 0u, (int32_t)k, v};
  }
}
__device__ inline bool fundamental_hit(float v, uint32_t k, const BoundOut& bo) {
  return v > bo.thr && (int32_t)k >= bo.start0 && (int32_t)k < bo.end0;
}

template <int Q>
__global__ void __launch_bounds__(kB) k_spec_post(const float2* __restrict__ Z, float* __restrict__ P, uint64_t M,
                                                  int64_t pstride, RootTable rt, const TrialDesc* __restrict__ descs) {
  const uint32_t tr = blockIdx.y;
  const float2* z = Z + (int64_t)tr * (int64_t)M;
  float* p = P + (int64_t)tr * pstride;
  const float mean_n = descs[tr].mean_n;
  const float std_n = descs[tr].std_n;
  const uint32_t m = (uint32_t)M, half = m / 2;
  const int lane = (int)(threadIdx.x & 63u);
  const uint32_t waves = gridDim.x * (kB / 64);
  const uint32_t wave0 = blockIdx.x * (kB / 64) + (threadIdx.x >> 6);
  const uint32_t nwchunks = (half + 1 + kStep - 1) / kStep;  // waves of 63 lower outputs cover 0..half
  for (uint32_t c = wave0; c < nwchunks; c += waves * Q) {
    float2 zk[Q], zm[Q];
    int32_t kq[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const uint32_t cw = c + (uint32_t)q * waves;
      const int32_t k = (int32_t)(cw * kStep) + lane - 1;
      kq[q] = k;
      zk[q] = zm[q] = make_float2(0.f, 0.f);
      if (cw < nwchunks && k >= 0 && (uint32_t)k <= half) {
        zk[q] = z[k];
        zm[q] = z[k == 0 ? 0 : m - (uint32_t)k];
      }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
      const int32_t k = kq[q];
      const bool valid = (c + (uint32_t)q * waves) < nwchunks && k >= 0 && (uint32_t)k <= half;
      float2 xl = make_float2(0.f, 0.f), xu = make_float2(0.f, 0.f);
      if (valid) {
        const float2 w = root(rt, (uint64_t)k);
        xl = untangle_one(zk[q], zm[q], w);
        // W_n^(M-k) = -conj(W_n^k); the midpoint is its own mirror and keeps
        // the lower formula, as the plain pass writes it
        xu = ((uint32_t)k == half) ? xl : untangle_one(zm[q], zk[q], make_float2(-w.x, w.y));
      }
      // X[k-1] from the lane below (zero below bin 0), X[M-k-1] from above
      const float2 xl_prev = shfl_f2(xl, lane > 0 ? lane - 1 : 0);
      const float2 xu_next = shfl_f2(xu, lane < 63 ? lane + 1 : 63);
      if (valid) {
        if (lane > 0) p[k] = normalised(xl, xl_prev, mean_n, std_n);
        if (lane < 63 && (uint32_t)k < half) p[m - (uint32_t)k] = normalised(xu, xu_next, mean_n, std_n);
      }
    }
  }
}

// The same spectrum pass over the transposed half-length spectrum that a
// last FFT pass writing contiguous rows leaves: row r (R rows of L
// elements) holds Z(r, e) = Z[r + R*e].  Bin k = r + R*e pairs with M - k,
// which is (R - r, L-1-e) for r >= 1 and, for r = 0, (0, L-e), i.e. the
// virtual row R = row 0 shifted by one column (Z[M] = Z[0]).  Interbinning
// takes X[k-1] = X(r-1, e) from the row below, in the same column.
//
// A workgroup owns lower rows [ra, ra+64) (1 <= r < R/2) and their mirror
// rows R - r, over 64 columns.  Wave w sweeps pair rows q = r0-1 .. r0+16
// (r0 = ra + 16w), one column per lane: it loads Z(q, e) (row contiguous)
// and the mirror Z(R-q, L-1-e) (row descending), untangles both bins with
// the plain pass's arithmetic, and keeps the previous pair row in registers
// for the interbin neighbours: P(q, e) uses X(q-1, e), and P(R-q', L-1-e)
// uses X(R-q'-1, L-1-e), the mirror of pair row q'+1.  Results are
// transposed through LDS and written as 64-bin runs of P (natural order).
// Rows 0 and R/2 and bin M pair within their own row: the grid's last row
// of workgroups computes them (spec_edge_bin), beside the tiles.
constexpr int kRowTile = 64;  // lower rows per workgroup (16 per wave)
constexpr int kColTile = 64;  // columns per workgroup (one per lane)
// the write-out gives each lane one row of the tile
static_assert(kRowTile == 64 && kColTile == 64, "k_spec_rows writes one tile row per lane");

// X[k] of the pair (Z[k], Z[M-k]) exactly as untangled_bin computes it:
// the root of min(k, M-k), mirrored for the upper bin.
// hi (optional): the root table's high factor of this root, when the
// caller already holds it; the product is the same cmul either way.
__device__ inline void untangle_pair(float2 zk, float2 zm, uint32_t k, uint32_t M, const RootTable& rt, float2& xk,
                                     float2& xm, const float2* hi = nullptr) {
  const bool upper = 2 * k > M;
  const uint32_t ri = upper ? M - k : k;
  const float2 w = hi ? cmul(*hi, ldg2(rt.lo, ri & rt.mask)) : root32(rt, ri);
  const float2 wm = make_float2(-w.x, w.y);
  xk = untangle_one(zk, zm, upper ? wm : w);
  xm = untangle_one(zm, zk, upper ? w : wm);
}

// Bins of rows 0 and R/2 (k = R*e and R/2 + R*e) and the Nyquist bin M of
// the transposed layout, one thread each, through the natural-index view.
// rshift: log2(R) when R is a power of two (the row and column of j by a
// mask and a shift), else 32 (by division).
__device__ inline float2 z_rows(const float2* y, uint32_t R, uint32_t L, uint32_t j, uint32_t rshift = 32) {
  const uint32_t M = R * L;
  if (j >= M) j -= M;
  return rshift < 32 ? y[(j & (R - 1)) * L + (j >> rshift)] : y[(j % R) * L + j / R];
}
__device__ inline float2 x_rows(const float2* y, uint32_t R, uint32_t L, uint32_t k, const RootTable& rt,
                                uint32_t rshift = 32) {
  const uint32_t M = R * L;
  float2 xk, xm;
  untangle_pair(z_rows(y, R, L, k, rshift), z_rows(y, R, L, M - k, rshift), k, M, rt, xk, xm);
  return xk;
}
// t < L: row 0; t < 2L: row R/2; t == 2L: bin M.  Returns the bin and its
// value; with bo, also the bound of the bin pair (k, k+1), whose second bin
// (row 1 or R/2 + 1, a tile's) is recomputed here through the same untangle
// and interbin arithmetic, so it equals the tile's value bit for bit.
template <bool BOUND>
__device__ inline float spec_edge_bin(const float2* y, float* p, uint32_t R, uint32_t L, uint32_t t,
                                      const RootTable& rt, float mean_n, float std_n, uint16_t* q, uint32_t& k) {
  const uint32_t M = R * L;
  k = t < L ? R * t : (t < 2 * L ? R / 2 + R * (t - L) : M);
  float2 x, prev = make_float2(0.f, 0.f);
  if (k == M) {
    // X[M] is the upper bin of pair 0
    float2 x0;
    untangle_pair(y[0], y[0], 0, M, rt, x0, x);
  } else {
    x = x_rows(y, R, L, k, rt);
  }
  if (k > 0) prev = x_rows(y, R, L, k - 1, rt);
  const float v = normalised(x, prev, mean_n, std_n);
  if constexpr (!BOUND) p[k] = v;
  if constexpr (BOUND) {
    float m = v;  // bin M has no partner (n/2 + 1 bins)
    if (k < M) m = fmaxf(v, normalised(x_rows(y, R, L, k + 1, rt), x, mean_n, std_n));
    q[k / 2] = bf16_up(m);
  }
  return v;
}

// BOUND: the bound spectrum and the fundamental's candidates instead of the
// spectra (the bound-filtered harmonic pass recomputes the few bins it sums
// exactly from Y); the tile's values are then upper bounds (normalised_up),
// and a value above the threshold is recomputed exactly for the fundamental.
template <bool HI_FIXED, bool BOUND>
__global__ void __launch_bounds__(kB) k_spec_rows(const float2* __restrict__ Y, float* __restrict__ P, uint32_t R,
                                                  uint32_t L, int64_t pstride, RootTable rt,
                                                  const TrialDesc* __restrict__ descs, BoundOut bo) {
  __shared__ float lo_t[kColTile][kRowTile + 1];
  __shared__ float up_t[kColTile][kRowTile + 1];
  __shared__ float2 edge_lo[kB / 64][kColTile];  // X of each wave's last lower row
  __shared__ float2 edge_hi[kB / 64][kColTile];  // X of each wave's first upper row
  const uint32_t tr = blockIdx.z;
  const uint32_t M = R * L, half_rows = R / 2;
  const float2* y = Y + (uint64_t)tr * M;
  float* p = P + (int64_t)tr * pstride;
  const float mean_n = descs[tr].mean_n, std_n = descs[tr].std_n;
  uint16_t* q = BOUND ? bo.Q + (int64_t)tr * bo.qstride : nullptr;
  auto tile_value = [&](float2 xv, float2 pv) {
    if constexpr (BOUND)
      return normalised_up(xv, pv, mean_n, std_n);
    else
      return normalised(xv, pv, mean_n, std_n);
  };
  if (blockIdx.y == gridDim.y - 1) {
    // the edge bins: L/64 workgroups of kB threads cover the 2L+1 of them
    const uint32_t t = blockIdx.x * kB + threadIdx.x;
    const bool valid = t <= 2 * L;
    uint32_t k = 0;
    float v = 0.f;
    if (valid) v = spec_edge_bin<BOUND>(y, p, R, L, t, rt, mean_n, std_n, q, k);
    if constexpr (BOUND) append_fundamental(valid && fundamental_hit(v, k, bo), descs[tr].slot, k, v, bo);
    return;
  }
  const uint32_t ra = 1 + blockIdx.y * kRowTile, ea = blockIdx.x * kColTile;
  const unsigned lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
  const uint32_t e = ea + lane, emir = L - 1 - e;
  constexpr int kWaveRows = kRowTile / (kB / 64);
  const uint32_t r0 = ra + wv * kWaveRows;
  // pair row q -> (X(q, e), X(R-q, L-1-e)); rows past R/2 are never used
  // and load row R/2 instead (in bounds, results discarded)
  auto pair_row = [&](uint32_t qq, float2& xlo, float2& xhi) {
    const uint32_t q = min(qq, half_rows);
    const uint32_t ih = q == 0 ? (e == 0 ? 0 : L - e) : (R - q) * L + emir;
    untangle_pair(ldg2(y, q * L + e), ldg2(y, ih), q + R * e, M, rt, xlo, xhi);
  };
  // HI_FIXED (R a power of two no larger than the root table's low span
  // 2^shift, checked by the launcher): for rows q in [1, R/2] of one column
  // the root index is q + R*e or its mirror R*(L-e) - q, whose high part
  // (index >> shift) is then the same for all of them, so the sweep loads
  // that factor once per lane.
  float2 hi_e = make_float2(0.f, 0.f);
  if constexpr (HI_FIXED) {
    const uint32_t k0 = min(r0, half_rows) + R * e;
    hi_e = ldg2(rt.hi, (2 * k0 > M ? M - k0 : k0) >> rt.shift);
  }
  auto sweep_row = [&](uint32_t qq, float2& xlo, float2& xhi) {
    const uint32_t q = min(qq, half_rows);  // >= 1 in the sweep
    untangle_pair(ldg2(y, q * L + e), ldg2(y, (R - q) * L + emir), q + R * e, M, rt, xlo, xhi,
                  HI_FIXED ? &hi_e : nullptr);
  };
  // results of rows outside [1, R/2) go to the spare slot kRowTile
  auto slot = [&](uint32_t qq) { return qq < half_rows ? qq - ra : (uint32_t)kRowTile; };
  // Branch-free sweep over the wave's own rows (so the loads of later rows
  // issue early); the interbin neighbours across wave boundaries come from
  // the adjacent waves through LDS, and only the workgroup's outer halo
  // rows (ra-1 below, ra+64 above) are computed twice.
  float2 xlo_prev = make_float2(0.f, 0.f), xhi_prev = make_float2(0.f, 0.f);
  float2 xlo_first = xlo_prev, xhi_last = xhi_prev;
  // With the high root factor fixed, every row's loads (both bins and the
  // low root factor) are issued before any is used: 16 rows of loads in
  // flight per wave (127 VGPRs, the occupancy the LDS tiles allow anyway)
  // instead of the few the compiler keeps when loads and arithmetic
  // interleave.  The arithmetic is sweep_row's.
  float2 zl[kWaveRows], zh[kWaveRows], wl[kWaveRows];
  if constexpr (HI_FIXED) {
#pragma unroll
    for (int i = 0; i < kWaveRows; ++i) {
      const uint32_t q = min(r0 + (uint32_t)i, half_rows);
      const uint32_t k = q + R * e;
      zl[i] = ldg2(y, q * L + e);
      zh[i] = ldg2(y, (R - q) * L + emir);
      wl[i] = ldg2(rt.lo, (2 * k > M ? M - k : k) & rt.mask);
    }
  }
#pragma unroll
  for (int i = 0; i < kWaveRows; ++i) {
    const uint32_t qq = r0 + (uint32_t)i;
    float2 xlo, xhi;
    if constexpr (HI_FIXED) {
      const uint32_t k = min(qq, half_rows) + R * e;
      const bool upper = 2 * k > M;
      const float2 w = cmul(hi_e, wl[i]);
      const float2 wm = make_float2(-w.x, w.y);
      xlo = untangle_one(zl[i], zh[i], upper ? wm : w);
      xhi = untangle_one(zh[i], zl[i], upper ? w : wm);
    } else {
      sweep_row(qq, xlo, xhi);
    }
    if (i == 0) {
      xlo_first = xlo;
      edge_hi[wv][lane] = xhi;
    } else {
      lo_t[lane][slot(qq)] = tile_value(xlo, xlo_prev);       // P(q)
      up_t[lane][slot(qq - 1)] = tile_value(xhi_prev, xhi);   // P(R-(q-1))
    }
    xlo_prev = xlo;
    xhi_prev = xhi;
  }
  edge_lo[wv][lane] = xlo_prev;
  xhi_last = xhi_prev;
  float2 below = make_float2(0.f, 0.f), above = make_float2(0.f, 0.f), next_lo;
  if (wv == 0) pair_row(ra - 1, below, next_lo);
  if (wv == kB / 64 - 1) pair_row(ra + kRowTile, next_lo, above);
  __syncthreads();
  if (wv > 0) below = edge_lo[wv - 1][lane];
  if (wv < kB / 64 - 1) above = edge_hi[wv + 1][lane];
  lo_t[lane][slot(r0)] = tile_value(xlo_first, below);
  up_t[lane][slot(r0 + kWaveRows - 1)] = tile_value(xhi_last, above);
  if constexpr (BOUND) {
    // P of the next tile's first row (ra + 64), the partner of this tile's
    // last row in a bound pair: its X was loaded for the upper halo, and the
    // interbin neighbour is this tile's last row.  Slot kRowTile is free in
    // every tile but the last, which has no such pair.
    if (wv == kB / 64 - 1 && ra + kRowTile < half_rows)
      lo_t[lane][kRowTile] = tile_value(next_lo, xlo_prev);
  }
  __syncthreads();
  const uint32_t rr = threadIdx.x & 63u;
  const uint32_t r = ra + rr;
  if constexpr (!BOUND) {
    if (r < half_rows) {
#pragma unroll 4
      for (uint32_t j = wv; j < (uint32_t)kColTile; j += kB / 64) {
        stg1(p, r + R * (ea + j), lo_t[j][rr]);
        stg1(p, (R - r) + R * (L - 1 - (ea + j)), up_t[j][rr]);
      }
    }
  } else {
    // Bound pairs (2u, 2u+1): rows 0/1 and R/2 / R/2+1 of every column are the
    // edge workgroups'; a tile's column holds 32 lower pairs, rows
    // (ra+2i+1, ra+2i+2), i < 32 (slot kRowTile for the last one's second
    // row), and 32 upper pairs, rows R-r for r = ra+63-2i and ra+62-2i.
    // Lane i < 32 bounds lower pair i, lane 32+i upper pair i: one 16-bit
    // store covers both runs of 32 entries.  Every value written is also
    // tested as a fundamental.
    const bool rv = r < half_rows;
    const bool lower = rr < 32;
    const uint32_t i = rr & 31u;
    const uint32_t s0 = lower ? 2 * i + 1 : 63 - 2 * i;  // slot of the pair's even row
    const uint32_t s1 = lower ? s0 + 1 : s0 - 1;         // and of its odd row
    const uint32_t rpair = ra + s0;                      // lower-half row of the even-row slot
    const bool pv = rpair < half_rows;
    bool any = false;
#pragma unroll 4
    for (uint32_t j = wv; j < (uint32_t)kColTile; j += kB / 64) {
      const float vl = lo_t[j][rr], vu = up_t[j][rr];
      if (rv) any |= fmaxf(vl, vu) > bo.thr;
      if (pv) {
        const float m = lower ? fmaxf(lo_t[j][s0], lo_t[j][s1]) : fmaxf(up_t[j][s0], up_t[j][s1]);
        const uint32_t k2 = lower ? rpair + R * (ea + j) : (R - rpair) + R * (L - 1 - (ea + j));
        q[k2 / 2] = bf16_up(m);
      }
    }
    // bounds above the threshold are rare: a second sweep over the tile, in
    // the waves that have any, recomputes those bins exactly (the same
    // untangle, interbin and normalisation of the same inputs as the exact
    // pass) and appends the ones above threshold
    if (__ballot(any) != 0) {
      const uint32_t slot_id = descs[tr].slot;
      const uint32_t rshift = HI_FIXED ? (uint32_t)__builtin_ctz(R) : 32u;
      auto exact = [&](uint32_t k) {
        const float2 prev = k > 0 ? x_rows(y, R, L, k - 1, rt, rshift) : make_float2(0.f, 0.f);
        return normalised(x_rows(y, R, L, k, rt, rshift), prev, mean_n, std_n);
      };
      for (uint32_t j = wv; j < (uint32_t)kColTile; j += kB / 64) {
        const uint32_t kl = r + R * (ea + j);
        const uint32_t ku = (R - r) + R * (L - 1 - (ea + j));
        float vl = lo_t[j][rr], vu = up_t[j][rr];
        const bool cl = rv && fundamental_hit(vl, kl, bo);
        const bool cu = rv && fundamental_hit(vu, ku, bo);
        if (cl) vl = exact(kl);
        if (cu) vu = exact(ku);
        append_fundamental(cl && vl > bo.thr, slot_id, kl, vl, bo);
        append_fundamental(cu && vu > bo.thr, slot_id, ku, vu, bo);
      }
    }
  }
}

// harmonic_sum_kernel (src/kernels.cu:34-100) + device_find_peaks
// (src/kernels.cu:391-423) for the fundamental and every fold, each over
// its own PeakFinder window (peakfinder.hpp:78-83).  Survivors are appended
// to `out` with one atomic per bin row of a wave (all folds together); the
// host orders them by bin.
//
// A wave covers kHarmWaveBins consecutive bins; lane l holds the D bins
// base + l + 64d.  Fold k's harmonic j of bin i is bin (j*i + 2^(k-1)) >> k,
// and with i = i0 + 64d that is ((j*i0 + 2^(k-1)) >> k) + d*64j/2^k exactly
// (2^k divides 64), so each harmonic costs one index per lane and D loads at
// constant instruction offsets from it: the gathers stay coalesced across
// lanes and the index arithmetic of harmonic_sums32 drops out of the loop.
// The running sums take the same addends in the same order as
// harmonic_sums32, so the thresholded values are bit-identical to it.  The
// loop only decides whether any value of a lane crosses the threshold; a
// wave with a crossing recomputes its bins one by one with harmonic_sums32,
// applies the fold windows and appends.
//
// Harmonics at 1/16 of the bin or less (step <= 4) touch only D*step + 1
// consecutive bins per wave, fewer than 64: the wave loads them once (one
// load per lane) and each d takes its value from the owning lane with
// ds_bpermute, instead of D loads of a few distinct floats each (2% faster
// harmonic pass).  The values are the same bins, so the sums are unchanged.
template <int K, int J, int D>
__device__ inline void add_harmonic(const PSX_GLOBAL float* p, uint32_t i0, float (&val)[D]) {
  constexpr uint32_t step = (64u * J) >> K;
  constexpr uint32_t c = 1u << (K - 1);
  if constexpr (step <= 4 && D * step + 1 < 64) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t s0 = (J * (i0 - lane) + c) >> K;  // the wave's first bin of this harmonic
    const uint32_t r = ((J * i0 + c) >> K) - s0;     // this lane's offset, <= step
    const float chunk = p[s0 + lane];
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const int addr = (int)((r + (uint32_t)d * step) << 2);
      val[d] += __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(chunk)));
    }
  } else {
    const PSX_GLOBAL float* q = p + ((J * i0 + c) >> K);
#pragma unroll
    for (int d = 0; d < D; ++d) val[d] += q[d * step];
  }
}
// odd harmonics J, J+2, ... below JEND of fold K, in increasing order
template <int K, int J, int JEND, int D>
__device__ inline void add_harmonics(const PSX_GLOBAL float* p, uint32_t i0, float (&val)[D]) {
  add_harmonic<K, J, D>(p, i0, val);
  if constexpr (J + 2 < JEND) add_harmonics<K, J + 2, JEND, D>(p, i0, val);
}
template <int K, int D>
__device__ inline bool fold_crosses(const float (&val)[D], float thr) {
  bool any = false;
#pragma unroll
  for (int d = 0; d < D; ++d) any |= fold_scale<K>(val[d]) > thr;
  return any;
}

template <int NF>
__global__ void __launch_bounds__(kB) k_harm_peaks(const float* __restrict__ P, int64_t pstride, Windows w,
                                                   const TrialDesc* __restrict__ descs, Cand16* __restrict__ out,
                                                   uint32_t cap, uint32_t* __restrict__ count) {
  constexpr int D = kHarmWaveBins / 64;
  const uint32_t tr = blockIdx.y;
  const float* p = P + (int64_t)tr * pstride;
  const PSX_GLOBAL float* pg = (const PSX_GLOBAL float*)p;
  const unsigned lane = threadIdx.x & 63u;
  const uint32_t hi = (uint32_t)w.hi;
  const uint32_t wstep = gridDim.x * (kB / 64) * kHarmWaveBins;
  for (uint32_t base = (uint32_t)w.lo + (blockIdx.x * (kB / 64) + (threadIdx.x >> 6)) * kHarmWaveBins; base < hi;
       base += wstep) {
    const uint32_t i0 = base + lane;
    float val[D];
    bool any = false;
    {
      const PSX_GLOBAL float* q = pg + i0;
#pragma unroll
      for (int d = 0; d < D; ++d) val[d] = q[64 * d];
#pragma unroll
      for (int d = 0; d < D; ++d) any |= val[d] > w.thr;
    }
    if constexpr (NF > 0) {
      add_harmonic<1, 1, D>(pg, i0, val);
      any |= fold_crosses<1, D>(val, w.thr);
    }
    if constexpr (NF > 1) {
      add_harmonic<2, 3, D>(pg, i0, val);  // 3/4 before 1/4, as the reference
      add_harmonic<2, 1, D>(pg, i0, val);
      any |= fold_crosses<2, D>(val, w.thr);
    }
    if constexpr (NF > 2) {
      add_harmonics<3, 1, 8, D>(pg, i0, val);
      any |= fold_crosses<3, D>(val, w.thr);
    }
    if constexpr (NF > 3) {
      add_harmonics<4, 1, 16, D>(pg, i0, val);
      any |= fold_crosses<4, D>(val, w.thr);
    }
    if constexpr (NF > 4) {
      add_harmonics<5, 1, 32, D>(pg, i0, val);
      any |= fold_crosses<5, D>(val, w.thr);
    }
    // survivors are rare: one ballot decides whether this wave appends at all
    if (__ballot(any) == 0) continue;
    const uint32_t slot = descs[tr].slot;
#pragma unroll 1
    for (int d = 0; d < D; ++d) {
      const uint32_t i = i0 + 64u * (uint32_t)d;
      const bool valid = i < hi;
      float s[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      if (valid) {
        float f[5];
        psx::harmonic_sums32<NF>(p, i, f);
        s[0] = p[i];
#pragma unroll
        for (int k = 0; k < NF; ++k) s[k + 1] = f[k];
      }
      // one atomic for the wave's hits of all folds at this d
      bool hit[NF + 1];
      uint64_t mask[NF + 1];
      uint32_t total = 0;
#pragma unroll
      for (int nh = 0; nh <= NF; ++nh) {
        hit[nh] = valid && (int32_t)i >= w.start[nh] && (int32_t)i < w.end[nh] && s[nh] > w.thr;
        mask[nh] = __ballot(hit[nh]);
        total += (uint32_t)__popcll(mask[nh]);
      }
      if (total == 0) continue;
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(count, total);
      b = __shfl(b, 0);
#pragma unroll
      for (int nh = 0; nh <= NF; ++nh) {
        if (hit[nh]) {
          const uint32_t pos = b + psx::mbcnt64(mask[nh]);
          if (pos < cap) out[pos] = Cand16{slot, (uint32_t)nh, (int32_t)i, s[nh]};
        }
        b += (uint32_t)__popcll(mask[nh]);
      }
    }
  }
}


// ------------------------------------------------------------------------
// Bound-filtered harmonic pass.
//
// The gather form above reads each spectrum once per fold span (8.5 spectra
// for four folds) and is bound by the L2-miss path: every XCD's 4 MB L2
// holds a fraction of one 16 MB spectrum, and the miss path moves about
// 7.5 TB/s whether the MALL or HBM serves it.  Candidates are rare, so this
// form first proves most bins empty from a smaller array:
//   Q[u] = max(P[2u], P[2u+1]) rounded up to bfloat16,
// one 16-bit entry per two bins (a quarter of the spectrum's bytes).  Every
// addend P[x] of a fold sum is at most Q[x/2], and float addition,
// fold_scale and the comparison are monotone (round to nearest never
// reverses an order), so the running sum of the Q entries taken in the
// reference's addend order is at least the exact running sum at every fold:
// where the bound of fold k stays at or below the threshold, so does fold
// k's exact value.  Only the bins whose bound crosses (about 1e-6 of
// noise bins, and every real detection) are summed exactly from P with
// harmonic_sums32, thresholded and appended, so the candidates are the
// gather form's bit for bit.  bfloat16 (8 significant bits) loosens the
// bound no more than half precision does at these values (the same rate of
// crossing noise bins in a float64 simulation), and rounds up with two
// integer operations where half precision needs a conversion, a check and
// a step.
//
// fold_scale<K>(v) > thr is tested as v > T[K], T[K] the largest float
// whose fold_scale is at most thr (found on the host by bisection over the
// float order; fold_scale is monotone), which keeps double arithmetic out
// of the loop.
//
// A wave covers kBoundWaveBins consecutive bins, lane l holding bins
// base + l + 64d (d < 8), as in the gather form.  Harmonic j/2^k of those
// 512 bins takes its addends from at most 512j/2^k + 1 consecutive source
// bins, one run of 256j/2^k + 1 entries of Q; the wave copies every run (and
// the fundamental's) into its own LDS area with global_load_lds_dwordx4
// (one instruction per run, 16 bytes per lane, no registers), waits once,
// and then reads each addend from LDS at a constant offset per d.  The
// wave's LDS area is private: no barriers.  Measured at 2^23 with 64 trials
// per launch: 0.36 ms against 0.80 ms for the gather form, the same
// candidates; writing the bounds costs the spectrum pass about 0.1 ms.

// Harmonic h (0-based, in the reference's addend order, fold 2 taking 3/4
// before 1/4) is j/2^k with k = hK(h), j = hJ(h); fold k ends at h = 2^k - 2.
__host__ __device__ constexpr int hK(int h) { return h < 1 ? 1 : h < 3 ? 2 : h < 7 ? 3 : h < 15 ? 4 : 5; }
__host__ __device__ constexpr int hJ(int h) {
  return h == 0 ? 1 : h == 1 ? 3 : h == 2 ? 1 : h < 7 ? 2 * (h - 3) + 1 : h < 15 ? 2 * (h - 7) + 1 : 2 * (h - 15) + 1;
}
constexpr int kBoundChunk = 8;  // entries per 16-byte chunk
constexpr int kBoundD = kBoundWaveBins / 64;
// chunks of harmonic h's run: its sources span at most j*W/2^k bins, W/2^k/2
// entries, plus one entry and the chunk alignment
__host__ __device__ constexpr int run_chunks(int h) {
  return ((hJ(h) * (int)kBoundWaveBins >> hK(h)) / (int)kBoundPool) / kBoundChunk + 2;
}
// LDS offset (entries) of run h; the fundamental's run (W/2 entries + the
// alignment) comes first, at 0
__host__ __device__ constexpr int run_offset(int h) {
  int r = ((int)kBoundWaveBins / (int)kBoundPool / kBoundChunk + 1) * kBoundChunk;
  for (int g = 0; g < h; ++g) r += run_chunks(g) * kBoundChunk;
  return r;
}
static_assert(run_chunks(30) <= 64 && (int)kBoundWaveBins / (int)kBoundPool / kBoundChunk + 1 <= 64,
              "one 64-lane instruction per run");

typedef float psx_f4v __attribute__((ext_vector_type(4)));

// The bound spectrum Q of every trial row, and the fundamental's threshold
// test (device_find_peaks over PeakFinder's nh = 0 window,
// src/kernels.cu:391-423, peakfinder.hpp:78-83) on the same read of P.
// Each lane takes four bins (one 16-byte load) and writes two entries.
__global__ void __launch_bounds__(kB) k_pool_peaks(const float* __restrict__ P, int64_t pstride, uint32_t nb,
                                                   Windows w, uint16_t* __restrict__ Q, int64_t qstride,
                                                   const TrialDesc* __restrict__ descs, Cand16* __restrict__ out,
                                                   uint32_t cap, uint32_t* __restrict__ count) {
  const uint32_t tr = blockIdx.y;
  const float* p = P + (int64_t)tr * pstride;
  uint16_t* q = Q + (int64_t)tr * qstride;
  const unsigned lane = threadIdx.x & 63u;
  const uint32_t start = (uint32_t)max(w.start[0], 0), end = (uint32_t)max(w.end[0], 0);
  const uint32_t ngroups = (nb + 3) / 4;
  const uint32_t stride = gridDim.x * kB;
  // whole-wave trip count, so the ballots below see every lane
  for (uint32_t g0 = blockIdx.x * kB + (threadIdx.x & ~63u); g0 < ngroups; g0 += stride) {
    const uint32_t g = g0 + lane;
    const uint32_t k0 = 4 * g;
    float x[4] = {-INFINITY, -INFINITY, -INFINITY, -INFINITY};
    if (g < ngroups) {
      // k0 + 3 < pstride: rows are padded to a multiple of 64 bins
      const psx_f4v v = __builtin_nontemporal_load((const PSX_GLOBAL psx_f4v*)(p + k0));
      x[0] = v.x;
      x[1] = k0 + 1 < nb ? v.y : -INFINITY;
      x[2] = k0 + 2 < nb ? v.z : -INFINITY;
      x[3] = k0 + 3 < nb ? v.w : -INFINITY;
      const uint32_t pair = (uint32_t)bf16_up(fmaxf(x[0], x[1])) | ((uint32_t)bf16_up(fmaxf(x[2], x[3])) << 16);
      *(PSX_GLOBAL uint32_t*)(q + 2 * g) = pair;
    }
    bool any = false;
#pragma unroll
    for (int j = 0; j < 4; ++j) any |= x[j] > w.thr;
    if (__ballot(any) == 0) continue;
    const uint32_t slot = descs[tr].slot;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t k = k0 + (uint32_t)j;
      const bool hit = x[j] > w.thr && k >= start && k < end;
      const uint64_t mask = __ballot(hit);
      if (mask == 0) continue;
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(count, (uint32_t)__popcll(mask));
      b = __shfl(b, 0);
      if (hit) {
        const uint32_t pos = b + psx::mbcnt64(mask);
        if (pos < cap) out[pos] = Cand16{slot, 0u, (int32_t)k, x[j]};
      }
    }
  }
}

struct FoldThr {
  float t[6];  // t[k]: fold k crosses iff its running sum > t[k]
};

// Copies one run of Q (entries s0..s1, rounded out to whole chunks) into
// the wave's LDS at dst; returns the first entry copied.
__device__ inline uint32_t stage_run(const PSX_GLOBAL uint16_t* q, uint32_t s0, uint32_t s1, uint16_t* dst,
                                     unsigned lane) {
  const uint32_t a0 = s0 & ~(uint32_t)(kBoundChunk - 1);
  const uint32_t n = (s1 - a0) / kBoundChunk + 1;
  if (lane < n)
    __builtin_amdgcn_global_load_lds((const PSX_GLOBAL void*)(q + a0 + kBoundChunk * lane),
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
  return a0;
}

template <int NH, int H>
struct BoundRuns {
  __device__ static void stage(const PSX_GLOBAL uint16_t* q, uint32_t base, uint16_t* my, uint32_t (&a0)[NH + 1],
                               unsigned lane) {
    if constexpr (H < NH) {
      constexpr int K = hK(H), J = hJ(H);
      constexpr uint32_t c = 1u << (K - 1);
      a0[H] = stage_run(q, ((J * base + c) >> K) / kBoundPool,
                        ((J * (base + kBoundWaveBins - 1) + c) >> K) / kBoundPool, my + run_offset(H), lane);
      BoundRuns<NH, H + 1>::stage(q, base, my, a0, lane);
    }
  }
  // adds harmonic H's bound to every running sum; at a fold's last addend
  // marks the bins whose bound crosses that fold's threshold
  __device__ static void add(const uint16_t* my, const uint32_t (&a0)[NH + 1], uint32_t i0, const FoldThr& ft,
                             float (&val)[kBoundD], bool (&fail)[kBoundD], float (&t)[kBoundD]) {
    if constexpr (H < NH) {
      constexpr int K = hK(H), J = hJ(H);
      constexpr uint32_t c = 1u << (K - 1);
      constexpr uint32_t step = (64u * J) >> K;  // source step between a lane's bins
      static_assert(step % kBoundPool == 0, "constant LDS offsets per d");
      const uint16_t* b = my + run_offset(H) + (((J * i0 + c) >> K) / kBoundPool - a0[H]);
      static_assert(kBoundD == 8, "lds_bf16x8 reads eight entries");
      lds_bf16x8<step / kBoundPool>(b, t);
#pragma unroll
      for (int d = 0; d < kBoundD; ++d) val[d] += t[d];
      if constexpr (((H + 2) & (H + 1)) == 0) {  // H = 2^K - 2: fold K complete
#pragma unroll
        for (int d = 0; d < kBoundD; ++d) fail[d] |= val[d] > ft.t[K];
      }
      BoundRuns<NH, H + 1>::add(my, a0, i0, ft, val, fail, t);
    }
  }
};

// Where the exact sums of the bins whose bound crosses take the spectrum
// from: P, or (rows != nullptr) bin by bin from the row-layout FFT output the
// spectrum pass read, through its untangle, interbin and normalisation
// (x_rows and normalised: for every bin the same operations on the same
// inputs as the spectrum pass, so the same value bit for bit).
struct ExactSrc {
  const float2* rows;  // Y of trial 0 (trial stride R*L), or nullptr: read P
  uint32_t R, L;
  uint32_t rshift;     // log2(R) for a power of two, else 32
  RootTable rt;
};
// harmonic_sums32's sums (the same addends in the same order) with the
// spectrum value of bin x given by f(x), one addend per iteration of a loop
// the compiler keeps rolled: this path runs for the rare bins whose bound
// crosses, and an unrolled recomputation of 2^NF spectrum values would hold
// enough registers to cost the main loop its occupancy.
template <int NF, class F>
__device__ inline void harmonic_sums_rolled(F f, uint32_t i, float (&s)[5]) {
  float val = f(i);
#pragma unroll 1
  for (int h = 0; h < (1 << NF) - 1; ++h) {
    const int K = hK(h), J = hJ(h);
    val += f(((uint32_t)J * i + (1u << (K - 1))) >> K);
    if (h == 0) s[0] = psx::fold_scale<1>(val);
    else if (h == 2) s[1] = psx::fold_scale<2>(val);
    else if (h == 6) s[2] = psx::fold_scale<3>(val);
    else if (h == 14) s[3] = psx::fold_scale<4>(val);
    else if (h == 30) s[4] = psx::fold_scale<5>(val);
  }
}

// Register budget: eight waves per SIMD (the LDS runs allow eight
// four-wave workgroups per CU) up to four folds; the rare exact path's
// scalar registers would otherwise take one of them.
constexpr int bound_waves(int nf) { return nf <= 4 ? 8 : 3; }

template <int NF, bool FROM_ROWS>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(bound_waves(NF)))) k_harm_bound(const float* __restrict__ P, int64_t pstride,
                                                   const uint16_t* __restrict__ Q, int64_t qstride, Windows w,
                                                   FoldThr ft, const TrialDesc* __restrict__ descs,
                                                   const ExactSrc* __restrict__ srcp,
                                                   uint32_t xcd_order, Cand16* __restrict__ out, uint32_t cap,
                                                   uint32_t* __restrict__ count) {
  constexpr int NH = (1 << NF) - 1;
  constexpr int WE = run_offset(NH);  // LDS entries per wave
  __shared__ uint16_t lds[WE * (kB / 64)];
  uint32_t tr = blockIdx.y, bx = blockIdx.x;
  if (xcd_order) {
    // workgroups reach the 8 XCDs round robin: hardware block b runs on XCD
    // b % 8, so trial t is given to XCD t % 8 whole (its bound spectrum
    // stays in one L2); the launcher asks for this only when 8 divides the
    // trial count
    const uint32_t b = blockIdx.y * gridDim.x + blockIdx.x;
    const uint32_t k = b >> 3;
    tr = (b & 7u) + 8u * (k / gridDim.x);
    bx = k % gridDim.x;
  }
  const float* p = P + (int64_t)tr * pstride;
  const PSX_GLOBAL uint16_t* qg = (const PSX_GLOBAL uint16_t*)(Q + (int64_t)tr * qstride);
  const unsigned lane = threadIdx.x & 63u;
  const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  uint16_t* my = lds + wave * WE;
  const uint32_t hi = (uint32_t)w.hi;
  const uint32_t wstep = gridDim.x * (kB / 64) * kBoundWaveBins;
  // registers the bound entries are read into, high halves only: their low
  // halves stay zero for the whole kernel
  float t[kBoundD];
#pragma unroll
  for (int d = 0; d < kBoundD; ++d) t[d] = 0.f;
  for (uint32_t base = (uint32_t)w.lo + (bx * (kB / 64) + wave) * kBoundWaveBins; base < hi; base += wstep) {
    const uint32_t i0 = base + lane;
    // the previous tile's LDS reads are complete before the area is refilled
    __asm__ volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
    uint32_t a0[NH + 1];
    a0[NH] = stage_run(qg, base / kBoundPool, (base + kBoundWaveBins - 1) / kBoundPool, my, lane);
    BoundRuns<NH, 0>::stage(qg, base, my, a0, lane);
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    float val[kBoundD];
    bool fail[kBoundD];
    lds_bf16x8<64 / kBoundPool>(my + (i0 / kBoundPool - a0[NH]), t);
#pragma unroll
    for (int d = 0; d < kBoundD; ++d) {
      val[d] = t[d];
      fail[d] = false;
    }
    BoundRuns<NH, 0>::add(my, a0, i0, ft, val, fail, t);
    bool any = false;
#pragma unroll
    for (int d = 0; d < kBoundD; ++d) any |= fail[d];
    if (__ballot(any) == 0) continue;
    // exact folds of the bins whose bound crossed (the fundamental is the
    // pooling pass's)
    const uint32_t slot = descs[tr].slot;
#pragma unroll
    for (int d = 0; d < kBoundD; ++d) {
      if (__ballot(fail[d]) == 0) continue;
      const uint32_t i = i0 + 64u * (uint32_t)d;
      const bool valid = fail[d] && i < hi;
      float s[5] = {0.f, 0.f, 0.f, 0.f, 0.f};
      if (valid) {
        if constexpr (FROM_ROWS) {
          // read from memory here, not held in scalar registers by the loop
          const ExactSrc src = *srcp;
          const uint32_t M = src.R * src.L;
          const float2* y = src.rows + (uint64_t)tr * M;
          const float mean_n = descs[tr].mean_n, std_n = descs[tr].std_n;
          auto pval = [&](uint32_t x) {
            const float2 xk = x_rows(y, src.R, src.L, x, src.rt, src.rshift);
            const float2 prev =
                x > 0 ? x_rows(y, src.R, src.L, x - 1, src.rt, src.rshift) : make_float2(0.f, 0.f);
            return normalised(xk, prev, mean_n, std_n);
          };
          harmonic_sums_rolled<NF>(pval, i, s);
        } else {
          psx::harmonic_sums32<NF>(p, i, s);
        }
      }
      bool hit[NF];
      uint64_t mask[NF];
      uint32_t total = 0;
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        hit[k] = valid && (int32_t)i >= w.start[k + 1] && (int32_t)i < w.end[k + 1] && s[k] > w.thr;
        mask[k] = __ballot(hit[k]);
        total += (uint32_t)__popcll(mask[k]);
      }
      if (total == 0) continue;
      uint32_t b = 0;
      if (lane == 0) b = atomicAdd(count, total);
      b = __shfl(b, 0);
#pragma unroll
      for (int k = 0; k < NF; ++k) {
        if (hit[k]) {
          const uint32_t pos = b + psx::mbcnt64(mask[k]);
          if (pos < cap) out[pos] = Cand16{slot, (uint32_t)(k + 1), (int32_t)i, s[k]};
        }
        b += (uint32_t)__popcll(mask[k]);
      }
    }
  }
}

// Largest float v with fold_scale<K>(v) <= thr, by bisection over the float
// order (fold_scale is monotone non-decreasing).
template <int K>
float fold_threshold(float thr) {
  auto from_order = [](int64_t o) {
    const uint32_t u = o >= 0 ? (uint32_t)o : ((uint32_t)(-o) | 0x80000000u);
    float f;
    std::memcpy(&f, &u, 4);
    return f;
  };
  int64_t lo = -(int64_t)0x7F7FFFFF, hi = 0x7F7FFFFF;  // -FLT_MAX .. FLT_MAX
  if (!(psx::fold_scale<K>(from_order(lo)) <= thr)) return -INFINITY;
  if (psx::fold_scale<K>(from_order(hi)) <= thr) return INFINITY;
  while (hi - lo > 1) {
    const int64_t mid = lo + (hi - lo) / 2;
    if (psx::fold_scale<K>(from_order(mid)) <= thr)
      lo = mid;
    else
      hi = mid;
  }
  return from_order(lo);
}

__global__ void k_zero_u32(uint32_t* p) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0u;
}

}  // namespace

int launch_spec_post(const float2* Z, float* P, uint64_t M, int64_t pstride, const RootTable& rt,
                     const TrialDesc* d_descs, uint32_t ntr, hipStream_t s) {
  if (M < 8 || M > 0x7FFFFFFFull) return psx::fail(PSX_ERR_ARG, "spec_post: unsupported length");
  constexpr int Q = 4;
  const uint64_t chunks = (M / 2 + 1 + kStep - 1) / kStep;  // one wave each
  const unsigned gx = psx::grid_for(chunks * 64, kB * Q, 2048);
  k_spec_post<Q><<<dim3(gx, ntr), kB, 0, s>>>(Z, P, M, pstride, rt, d_descs);
  PSX_CHECK_LAUNCH("k_spec_post");
  return PSX_OK;
}

int launch_spec_rows(const float2* Y, float* P, uint32_t R, uint32_t L, int64_t pstride, const RootTable& rt,
                     const TrialDesc* d_descs, uint32_t ntr, hipStream_t s, const SpecBound* bound) {
  if (R < 4 || R % 2 || L % kColTile || (uint64_t)R * L > 0x7FFFFFFFull || pstride < (int64_t)R * L + 1)
    return psx::fail(PSX_ERR_ARG, "spec_rows: unsupported layout %u x %u", R, L);
  // the bound pairs of a tile's rows need an even tile base row minus one
  if (bound && (R % 4 || bound->qstride < bound_stride(pstride)))
    return psx::fail(PSX_ERR_ARG, "spec_rows: unsupported bound layout %u x %u", R, L);
  // row tiles of rows [1, R/2), then one row of workgroups for the edge bins
  const unsigned row_tiles = (R / 2 - 1 + kRowTile - 1) / kRowTile;
  static_assert(kColTile * 2 <= kB, "the edge row needs 2L+1 threads from L/kColTile workgroups");
  const dim3 grid(L / kColTile, row_tiles + 1, ntr);
  BoundOut bo{};
  if (bound)
    bo = BoundOut{bound->Q, bound->qstride, max(bound->w.start[0], 0), max(bound->w.end[0], 0),
                  bound->w.thr, bound->out, bound->cap, bound->count};
  const bool hi_fixed = (R & (R - 1)) == 0 && R <= (1u << rt.shift);
  if (!bound && !P) return psx::fail(PSX_ERR_ARG, "spec_rows: no spectrum buffer");
#define PSX_SPEC_ROWS(HF, BD) k_spec_rows<HF, BD><<<grid, kB, 0, s>>>(Y, P, R, L, pstride, rt, d_descs, bo)
  if (!bound)
    hi_fixed ? PSX_SPEC_ROWS(true, false) : PSX_SPEC_ROWS(false, false);
  else
    hi_fixed ? PSX_SPEC_ROWS(true, true) : PSX_SPEC_ROWS(false, true);
#undef PSX_SPEC_ROWS
  PSX_CHECK_LAUNCH("k_spec_rows");
  return PSX_OK;
}

int launch_harm_peaks(const float* P, int64_t pstride, const Windows& w, uint32_t nfolds, const TrialDesc* d_descs,
                      uint32_t ntr, Cand16* out, uint32_t cap, uint32_t* count, hipStream_t s) {
  if (w.hi <= w.lo) return PSX_OK;
  // 32-bit bin arithmetic: j*i < 32 * 2^27
  if (w.hi > (1ll << 27)) return psx::fail(PSX_ERR_UNSUPPORTED, "harm_peaks: more than 2^27 bins");
  // a wave reads up to kHarmWaveBins - 1 bins past w.hi: the spectrum
  // buffer holds that many floats beyond the last trial's row
  if (w.lo < 0 || w.hi > pstride) return psx::fail(PSX_ERR_ARG, "harm_peaks: window outside the spectrum row");
  const unsigned gx = psx::grid_for((uint64_t)(w.hi - w.lo), (uint64_t)kHarmWaveBins * (kB / 64), 2048);
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


float fold_threshold_of(int k, float thr) {
  switch (k) {
    case 1: return fold_threshold<1>(thr);
    case 2: return fold_threshold<2>(thr);
    case 3: return fold_threshold<3>(thr);
    case 4: return fold_threshold<4>(thr);
    case 5: return fold_threshold<5>(thr);
    default: return INFINITY;
  }
}

int64_t bound_stride(int64_t pstride) {
  // a wave's staging reads up to kBoundWaveBins - 1 bins past the window
  // end, plus a chunk of alignment: per-row slack covers both
  return (pstride + kBoundWaveBins) / kBoundPool + 64;
}

int launch_pool_peaks(const float* P, int64_t pstride, uint64_t nb, const Windows& w, uint16_t* Q, int64_t qstride,
                      const TrialDesc* d_descs, uint32_t ntr, Cand16* out, uint32_t cap, uint32_t* count,
                      hipStream_t s) {
  if (nb < 1 || nb > (1ull << 27) || (int64_t)nb > pstride || pstride % 4 || qstride < bound_stride(pstride))
    return psx::fail(PSX_ERR_ARG, "pool_peaks: unsupported row layout");
  const unsigned gx = psx::grid_for((nb + 3) / 4, kB, 1024);
  k_pool_peaks<<<dim3(gx, ntr), kB, 0, s>>>(P, pstride, (uint32_t)nb, w, Q, qstride, d_descs, out, cap, count);
  PSX_CHECK_LAUNCH("k_pool_peaks");
  return PSX_OK;
}

int write_rows_source(const RowsSource& rs, void* d_src) {
  static_assert(sizeof(ExactSrc) <= kRowsSourceBytes, "kRowsSourceBytes");
  if (!d_src || !rs.Y || rs.R < 4 || (uint64_t)rs.R * rs.L > 0x7FFFFFFFull)
    return psx::fail(PSX_ERR_ARG, "write_rows_source: bad layout");
  const uint32_t R = rs.R;
  const ExactSrc e{rs.Y, R, rs.L, (R & (R - 1)) == 0 ? (uint32_t)__builtin_ctz(R) : 32u, rs.rt};
  PSX_CHECK_HIP(hipMemcpy(d_src, &e, sizeof(e), hipMemcpyHostToDevice));
  return PSX_OK;
}

int launch_harm_bound(const float* P, int64_t pstride, const uint16_t* Q, int64_t qstride, const Windows& w,
                      uint32_t nfolds, const TrialDesc* d_descs, uint32_t ntr, Cand16* out, uint32_t cap,
                      uint32_t* count, hipStream_t s, const RowsSource* rows, const void* d_rows_src) {
  if (w.hi <= w.lo) return PSX_OK;
  if (w.hi > (1ll << 27)) return psx::fail(PSX_ERR_UNSUPPORTED, "harm_bound: more than 2^27 bins");
  if (w.lo < 0 || w.hi > pstride || qstride < bound_stride(pstride))
    return psx::fail(PSX_ERR_ARG, "harm_bound: window outside the spectrum row");
  if (!rows && !P) return psx::fail(PSX_ERR_ARG, "harm_bound: no spectrum for the exact sums");
  if (rows && (!d_rows_src || w.hi > (int64_t)rows->R * rows->L + 1 ||
               (uint64_t)rows->R * rows->L > 0x7FFFFFFFull))
    return psx::fail(PSX_ERR_ARG, "harm_bound: window outside the row layout");
  const ExactSrc* src = rows ? (const ExactSrc*)d_rows_src : nullptr;
  FoldThr ft{};
  for (int k = 1; k <= 5; ++k) ft.t[k] = fold_threshold_of(k, w.thr);
  const unsigned gx = psx::grid_for((uint64_t)(w.hi - w.lo), (uint64_t)kBoundWaveBins * (kB / 64), 2048);
  const uint32_t xcd = ntr % 8 == 0 ? 1u : 0u;
  const dim3 g(gx, ntr);
  const uint16_t* q = Q;
#define PSX_HB(NF, FR) \
  k_harm_bound<NF, FR><<<g, kB, 0, s>>>(P, pstride, q, qstride, w, ft, d_descs, src, xcd, out, cap, count)
  const bool fr = rows != nullptr;
  switch (nfolds) {
    case 1: fr ? PSX_HB(1, true) : PSX_HB(1, false); break;
    case 2: fr ? PSX_HB(2, true) : PSX_HB(2, false); break;
    case 3: fr ? PSX_HB(3, true) : PSX_HB(3, false); break;
    case 4: fr ? PSX_HB(4, true) : PSX_HB(4, false); break;
    case 5: fr ? PSX_HB(5, true) : PSX_HB(5, false); break;
    default: return psx::fail(PSX_ERR_ARG, "nfolds must be 1..5");
  }
#undef PSX_HB
  PSX_CHECK_LAUNCH("k_harm_bound");
  return PSX_OK;
}

int launch_zero(uint32_t* p, hipStream_t s) {
  k_zero_u32<<<1, 64, 0, s>>>(p);
  PSX_CHECK_LAUNCH("k_zero_u32");
  return PSX_OK;
}

}  // namespace batch
}  // namespace psx
