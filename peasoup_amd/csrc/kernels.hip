// kernels.hip - per-stage kernels of the acceleration-search core for gfx950.
//
// One entry point per function of the reference's kernel layer
// (include/kernels/kernels.h); each kernel cites the reference kernel whose
// arithmetic it reproduces.  All elementwise stages are HBM-bound: they use
// 16-byte vector accesses where the layout allows and grid-stride loops sized
// to keep every CU busy.  Device code is compiled with -ffp-contract=off; the
// multiply-adds that the reference's nvcc build contracts are written as
// explicit fma so that results match the CPU restatement bit for bit.
#include <cmath>
#include <mutex>

#include "common.hpp"
#include "kernels_dev.hpp"
#include "reduce_async.hpp"
#include "trace.hpp"

using psx::as_stream;

namespace {

constexpr unsigned kBlock = 256;

// ---------------------------------------------------------------- resample
// resample_kernelII (src/kernels.cu:342-353): out[i] = in[rn(fma(i*af, i-n, i))].
__global__ void __launch_bounds__(kBlock) k_resample2(const float* __restrict__ in, float* __restrict__ out,
                                                      uint64_t n, double af) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t n4 = n / 4;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
    float4 r;
    r.x = in[psx::resample_index(4 * q + 0, n, af)];
    r.y = in[psx::resample_index(4 * q + 1, n, af)];
    r.z = in[psx::resample_index(4 * q + 2, n, af)];
    r.w = in[psx::resample_index(4 * q + 3, n, af)];
    reinterpret_cast<float4*>(out)[q] = r;
  }
  for (uint64_t i = 4 * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    out[i] = in[psx::resample_index(i, n, af)];
}

// resample_kernel (src/kernels.cu:327-339): id + af*((id-n/2)^2 - (n/2)^2),
// contracted as fma(af, fma(t, t, -(h*h)), id).
__global__ void __launch_bounds__(kBlock) k_resample1(const float* __restrict__ in, float* __restrict__ out,
                                                      uint64_t n, double af) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const double h = (double)n / 2.0;
  const double hh = h * h;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const double di = (double)i;
    const double t = di - h;
    const double q = fma(t, t, -hh);
    uint64_t j = (uint64_t)rint(fma(af, q, di));
    j = j < n ? j : n - 1;
    out[i] = in[j];
  }
}

// ---------------------------------------------------------- spectrum forming
// power_series_kernel (src/kernels.cu:216-234): |X|, 0 where |X|^2 == 0.
__global__ void __launch_bounds__(kBlock) k_amplitude(const float2* __restrict__ X, float* __restrict__ A,
                                                      uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
    const float2 x = X[i];
    const float z = fmaf(x.x, x.x, x.y * x.y);
    A[i] = (z == 0.0f) ? 0.0f : sqrtf(z);
  }
}

// bin_interbin_series_kernel (src/kernels.cu:238-259): two bins per thread,
// one 16-byte load for the pair plus the preceding bin.
__global__ void __launch_bounds__(kBlock) k_interbin(const float2* __restrict__ X, float* __restrict__ A,
                                                     uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t npair = nb / 2;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < npair; q += stride) {
    const float4 cur = reinterpret_cast<const float4*>(X)[q];
    const float2 prev = q > 0 ? X[2 * q - 1] : make_float2(0.f, 0.f);
    float2 o;
    o.x = psx::interbin(make_float2(cur.x, cur.y), prev);
    o.y = psx::interbin(make_float2(cur.z, cur.w), make_float2(cur.x, cur.y));
    reinterpret_cast<float2*>(A)[q] = o;
  }
  if (nb & 1) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      const uint64_t k = nb - 1;
      A[k] = psx::interbin(X[k], k > 0 ? X[k - 1] : make_float2(0.f, 0.f));
    }
  }
}

// ------------------------------------------------------------- dereddening
// median_scrunch5_kernel (src/kernels.cu:961-972).
__global__ void __launch_bounds__(kBlock) k_scrunch5(const float* __restrict__ in, float* __restrict__ out,
                                                     uint64_t oc) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < oc; i += stride) {
    const float* p = in + 5 * i;
    out[i] = psx::median5(p[0], p[1], p[2], p[3], p[4]);
  }
}

// median_scrunch5 special cases count = 1..4 (src/kernels.cu:981-997).
__global__ void k_scrunch_small(const float* __restrict__ in, float* __restrict__ out, int count) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  if (count == 1) out[0] = in[0];
  else if (count == 2) out[0] = 0.5f * (in[0] + in[1]);
  else if (count == 3) out[0] = psx::median3(in[0], in[1], in[2]);
  else out[0] = psx::median4(in[0], in[1], in[2], in[3]);
}

// linear_stretch_functor (src/kernels.cu:1010-1024).  step is computed on
// the host in float as the reference does.
__global__ void __launch_bounds__(kBlock) k_stretch(const float* __restrict__ in, uint64_t in_count,
                                                    float* __restrict__ out, uint64_t out_count, float step) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < out_count; i += stride)
    out[i] = psx::stretch_at(in, in_count, (unsigned)i, step);
}

// Dereddener::calculate_median zone copy (dereddener.hpp:53-61): writes the
// stretched median of zone z into median[lo, hi).
__global__ void __launch_bounds__(kBlock) k_stretch_zone(const float* __restrict__ in, uint64_t in_count,
                                                         float* __restrict__ out, uint64_t lo, uint64_t hi,
                                                         float step) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = lo + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < hi; i += stride)
    out[i] = psx::stretch_at(in, in_count, (unsigned)i, step);
}

// divide_c_by_f_kernel (src/kernels.cu:1040-1049).
__global__ void __launch_bounds__(kBlock) k_divide(float2* __restrict__ X, const float* __restrict__ f,
                                                   uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
    float2 x = X[i];
    if (i < 5) {
      x = make_float2(0.f, 0.f);
    } else {
      const float d = f[i];
      x.x = x.x / d;
      x.y = x.y / d;
    }
    X[i] = x;
  }
}

// zap_birdies_kernel (src/kernels.cu:1063-1084): bins
// [floor((f-w)/bw), ceil((f+w)/bw)) of each birdie, clamped to [0, nb-1),
// become 1+0i.  One block per birdie; the bin range is computed in float as
// the reference does and clamped before any integer conversion.
__global__ void __launch_bounds__(kBlock) k_zap(float2* __restrict__ X, const float* __restrict__ freqs,
                                                const float* __restrict__ widths, float bin_width, uint64_t nb) {
  const float f = freqs[blockIdx.x];
  const float w = widths[blockIdx.x];
  const double lo_f = (double)floorf((f - w) / bin_width);
  const double hi_f = (double)ceilf((f + w) / bin_width);
  if (!(lo_f < (double)nb)) return;  // also rejects NaN
  const uint64_t lo = lo_f > 0.0 ? (uint64_t)lo_f : 0;
  const uint64_t hi = hi_f >= (double)(nb - 1) ? nb - 1 : (hi_f > 0.0 ? (uint64_t)hi_f : 0);
  for (uint64_t i = lo + threadIdx.x; i < hi; i += blockDim.x) X[i] = make_float2(1.0f, 0.0f);
}

// --------------------------------------------------------------- reductions
// Deterministic two-level sum of x and of float(x*x) in double (GPU_mean /
// GPU_rms, src/kernels.cu:433-461, accumulate float terms in float).
constexpr unsigned kReduceBlocks = 1024;

// skip (optional): a stopped device baseline loop, whose later iterations
// need no sums.
__global__ void __launch_bounds__(kBlock) k_sum_sq(const float* __restrict__ x, uint64_t first, uint64_t n,
                                                   double2* __restrict__ partial,
                                                   const psx::BaselineState* __restrict__ skip) {
  if (skip && skip->done) return;
  double s1 = 0.0, s2 = 0.0;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = first + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
    const float v = x[i];
    const float sq = v * v;
    s1 += (double)v;
    s2 += (double)sq;
  }
  __shared__ double r1[kBlock], r2[kBlock];
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (unsigned o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r1[threadIdx.x] += r1[threadIdx.x + o];
      r2[threadIdx.x] += r2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) partial[blockIdx.x] = make_double2(r1[0], r2[0]);
}

__global__ void __launch_bounds__(kBlock) k_sum_final(const double2* __restrict__ partial, unsigned np,
                                                      double2* __restrict__ out,
                                                      const psx::BaselineState* __restrict__ skip) {
  if (skip && skip->done) return;
  double s1 = 0.0, s2 = 0.0;
  for (unsigned i = threadIdx.x; i < np; i += blockDim.x) {
    s1 += partial[i].x;
    s2 += partial[i].y;
  }
  __shared__ double r1[kBlock], r2[kBlock];
  r1[threadIdx.x] = s1;
  r2[threadIdx.x] = s2;
  __syncthreads();
  for (unsigned o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      r1[threadIdx.x] += r1[threadIdx.x + o];
      r2[threadIdx.x] += r2[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = make_double2(r1[0], r2[0]);
}

// normalisation_kernel (src/kernels.cu:496-507).
__global__ void __launch_bounds__(kBlock) k_normalise(float* __restrict__ p, uint64_t n, float mean, float sigma) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  const uint64_t n4 = n / 4;
  for (uint64_t q = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n4; q += stride) {
    float4 v = reinterpret_cast<float4*>(p)[q];
    v.x = (v.x - mean) / sigma;
    v.y = (v.y - mean) / sigma;
    v.z = (v.z - mean) / sigma;
    v.w = (v.w - mean) / sigma;
    reinterpret_cast<float4*>(p)[q] = v;
  }
  for (uint64_t i = 4 * n4 + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    p[i] = (p[i] - mean) / sigma;
}

// GPU_remove_baseline body (src/kernels.cu:471-473): x -= mean.
__global__ void __launch_bounds__(kBlock) k_subtract(float* __restrict__ x, uint64_t n, float mean) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] -= mean;
}

// Device-resident remove_baseline loop (the host loop of
// psx_remove_baseline with its arithmetic): each step subtracts the mean of
// the sums in `sums` unless the loop has already stopped, then records
// whether it continues.
__global__ void __launch_bounds__(kBlock) k_subtract_dev(float* __restrict__ x, uint64_t n,
                                                         const double2* __restrict__ sums,
                                                         const psx::BaselineState* __restrict__ st) {
  if (st->done) return;
  const float mean = (float)sums->x / (float)(int64_t)n;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) x[i] -= mean;
}
__global__ void k_baseline_step(uint64_t n, const double2* __restrict__ sums, psx::BaselineState* __restrict__ st) {
  if (blockIdx.x != 0 || threadIdx.x != 0 || st->done) return;
  const float mean = (float)sums->x / (float)(int64_t)n;
  // do { ...; if (count++ > 100) break; } while (fabsf(mean) > 5e-7 * (double)n);
  const int32_t count = st->count;
  st->count = count + 1;
  st->done = (count > 100 || !((double)fabsf(mean) > 5e-7 * (double)n)) ? 1 : 0;
}
__global__ void k_baseline_init(psx::BaselineState* st) {
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->count = 0;
    st->done = 0;
  }
}

// GPU_fill (src/kernels.cu:482-488).
__global__ void __launch_bounds__(kBlock) k_fill(float* __restrict__ x, uint64_t start, uint64_t end, float v) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = start + (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < end; i += stride) x[i] = v;
}

// conversion_kernel (src/kernels.cu:1171-1180).
template <class T>
__global__ void __launch_bounds__(kBlock) k_convert(const T* __restrict__ in, float* __restrict__ out, uint64_t n) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) out[i] = (float)in[i];
}

// ---------------------------------------------------------- harmonic summing
// harmonic_sum_kernel (src/kernels.cu:34-100): fold pointers are passed by
// value, so no pointer table crosses PCIe per call.
struct FoldPtrs {
  float* f[5];
};

template <int NF>
__global__ void __launch_bounds__(kBlock) k_harmonic(const float* __restrict__ P, FoldPtrs folds, uint64_t nb) {
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
    float s[5];
    psx::harmonic_sums<NF>(P, i, s);
#pragma unroll
    for (int k = 0; k < NF; ++k) folds.f[k][i] = s[k];
  }
}

// ------------------------------------------------------------- peak finding
// device_find_peaks (src/kernels.cu:391-423) as a two-kernel ordered
// compaction: per-tile counts, then each tile sums the counts of the tiles
// before it and writes its survivors in index order using wave ballots and
// mbcnt prefix sums.
constexpr unsigned kPeakTile = kBlock * 16;

__global__ void __launch_bounds__(kBlock) k_peak_count(const float* __restrict__ p, int64_t start, int64_t end,
                                                       float thr, uint32_t* __restrict__ counts) {
  const int64_t base = start + (int64_t)blockIdx.x * kPeakTile;
  unsigned c = 0;
#pragma unroll 4
  for (unsigned r = 0; r < kPeakTile / kBlock; ++r) {
    const int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
    if (i < end && p[i] > thr) ++c;
  }
  __shared__ unsigned red[kBlock];
  red[threadIdx.x] = c;
  __syncthreads();
  for (unsigned o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[blockIdx.x] = red[0];
}

__global__ void __launch_bounds__(kBlock) k_peak_write(const float* __restrict__ p, int64_t start, int64_t end,
                                                       float thr, const uint32_t* __restrict__ counts,
                                                       unsigned ntiles, int32_t* __restrict__ oidx,
                                                       float* __restrict__ osnr, uint32_t cap,
                                                       uint32_t* __restrict__ total) {
  __shared__ unsigned red[kBlock];
  __shared__ unsigned wave_cnt[kBlock / 64];
  // exclusive prefix of the tile counts before this tile
  unsigned pre = 0;
  for (unsigned t = threadIdx.x; t < blockIdx.x; t += kBlock) pre += counts[t];
  red[threadIdx.x] = pre;
  __syncthreads();
  for (unsigned o = kBlock / 2; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  unsigned offset = red[0];
  const unsigned mine = counts[blockIdx.x];
  if (blockIdx.x == ntiles - 1 && threadIdx.x == 0) *total = offset + mine;
  if (mine == 0) return;
  const unsigned lane = threadIdx.x & 63u;
  const unsigned wave = threadIdx.x >> 6;
  const int64_t base = start + (int64_t)blockIdx.x * kPeakTile;
  for (unsigned r = 0; r < kPeakTile / kBlock; ++r) {
    const int64_t i = base + (int64_t)r * kBlock + threadIdx.x;
    float v = 0.f;
    bool hit = false;
    if (i < end) {
      v = p[i];
      hit = v > thr;
    }
    const unsigned n_round = __syncthreads_count(hit);
    if (n_round == 0) continue;
    const uint64_t mask = __ballot(hit);
    const unsigned lane_pre = psx::mbcnt64(mask);
    if (lane == 0) wave_cnt[wave] = (unsigned)__popcll(mask);
    __syncthreads();
    unsigned wpre = 0;
    for (unsigned w = 0; w < wave; ++w) wpre += wave_cnt[w];
    if (hit) {
      const unsigned pos = offset + wpre + lane_pre;
      if (pos < cap) {
        oidx[pos] = (int32_t)i;
        osnr[pos] = v;
      }
    }
    offset += n_round;
    __syncthreads();
  }
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
  if (blockIdx.x == 0 && threadIdx.x == 0) *p = v;
}

inline float stretch_step(uint64_t in_count, uint64_t out_count) {
  // linear_stretch_functor ctor (src/kernels.cu:1015-1016): float / float
  return (float)(in_count - 1) / (float)(out_count - 1);
}

}  // namespace

// ============================================================ C ABI
extern "C" int psx_resample(const float* in, float* out, uint64_t n, float acc, float tsamp, psx_stream s) {
  PSX_REQUIRE(in && out && n > 0, "psx_resample: bad arguments");
  PSX_REQUIRE(in != out, "psx_resample: in and out may not alias");
  PSX_REQUIRE(((uintptr_t)out & 15) == 0, "psx_resample: out must be 16-byte aligned");
  const double af = psx::accel_fact(acc, tsamp);
  PSX_REQUIRE(fabs(af) * (double)n < 1.0, "psx_resample: |a*tsamp/2c| * n must be < 1 (got %g)",
              fabs(af) * (double)n);
  k_resample2<<<psx::grid_for(n / 4 + 1, kBlock), kBlock, 0, as_stream(s)>>>(in, out, n, af);
  PSX_CHECK_LAUNCH("k_resample2");
  return psx::ok();
}

extern "C" int psx_resample_v1(const float* in, float* out, uint64_t n, float acc, float tsamp, psx_stream s) {
  PSX_REQUIRE(in && out && n > 0, "psx_resample_v1: bad arguments");
  PSX_REQUIRE(in != out, "psx_resample_v1: in and out may not alias");
  const double af = psx::accel_fact(acc, tsamp);
  PSX_REQUIRE(fabs(af) * (double)n < 1.0, "psx_resample_v1: |a*tsamp/2c| * n must be < 1");
  k_resample1<<<psx::grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(in, out, n, af);
  PSX_CHECK_LAUNCH("k_resample1");
  return psx::ok();
}

extern "C" int psx_form_amplitude(const psx_c32* X, float* A, uint64_t nb, psx_stream s) {
  PSX_REQUIRE(X && A, "psx_form_amplitude: null pointer");
  if (!nb) return psx::ok();
  k_amplitude<<<psx::grid_for(nb, kBlock), kBlock, 0, as_stream(s)>>>((const float2*)X, A, nb);
  PSX_CHECK_LAUNCH("k_amplitude");
  return psx::ok();
}

extern "C" int psx_form_interbin(const psx_c32* X, float* A, uint64_t nb, psx_stream s) {
  PSX_REQUIRE(X && A, "psx_form_interbin: null pointer");
  PSX_REQUIRE(((uintptr_t)X & 15) == 0 && ((uintptr_t)A & 7) == 0, "psx_form_interbin: misaligned buffers");
  if (!nb) return psx::ok();
  k_interbin<<<psx::grid_for(nb / 2 + 1, kBlock), kBlock, 0, as_stream(s)>>>((const float2*)X, A, nb);
  PSX_CHECK_LAUNCH("k_interbin");
  return psx::ok();
}

extern "C" int psx_median_scrunch5(const float* in, uint64_t count, float* out, psx_stream s) {
  PSX_REQUIRE(in && out && count > 0, "psx_median_scrunch5: bad arguments");
  if (count < 5) {
    k_scrunch_small<<<1, 64, 0, as_stream(s)>>>(in, out, (int)count);
    PSX_CHECK_LAUNCH("k_scrunch_small");
    return psx::ok();
  }
  const uint64_t oc = count / 5;
  k_scrunch5<<<psx::grid_for(oc, kBlock), kBlock, 0, as_stream(s)>>>(in, out, oc);
  PSX_CHECK_LAUNCH("k_scrunch5");
  return psx::ok();
}

extern "C" int psx_linear_stretch(const float* in, uint64_t in_count, float* out, uint64_t out_count,
                                  psx_stream s) {
  PSX_REQUIRE(in && out && in_count > 0 && out_count > 1, "psx_linear_stretch: bad arguments");
  PSX_REQUIRE(out_count <= 0xFFFFFFFFull, "psx_linear_stretch: out_count exceeds 32-bit index");
  k_stretch<<<psx::grid_for(out_count, kBlock), kBlock, 0, as_stream(s)>>>(in, in_count, out, out_count,
                                                                           stretch_step(in_count, out_count));
  PSX_CHECK_LAUNCH("k_stretch");
  return psx::ok();
}

extern "C" int psx_divide_c_by_f(psx_c32* X, const float* f, uint64_t nb, psx_stream s) {
  PSX_REQUIRE(X && f, "psx_divide_c_by_f: null pointer");
  if (!nb) return psx::ok();
  k_divide<<<psx::grid_for(nb, kBlock), kBlock, 0, as_stream(s)>>>((float2*)X, f, nb);
  PSX_CHECK_LAUNCH("k_divide");
  return psx::ok();
}

extern "C" int psx_zap_birdies(psx_c32* X, const float* freqs, const float* widths, uint32_t nbirdies,
                               float bin_width, uint64_t nb, psx_stream s) {
  if (!nbirdies || !nb) return psx::ok();
  PSX_REQUIRE(X && freqs && widths && bin_width > 0, "psx_zap_birdies: bad arguments");
  PSX_REQUIRE(nbirdies <= 0x7FFFFFFFu, "psx_zap_birdies: too many birdies");
  k_zap<<<nbirdies, kBlock, 0, as_stream(s)>>>((float2*)X, freqs, widths, bin_width, nb);
  PSX_CHECK_LAUNCH("k_zap");
  return psx::ok();
}

extern "C" uint64_t psx_deredden_work_floats(uint64_t nb) { return nb / 5 + nb / 25 + nb / 125 + 3; }

// Dereddener::calculate_median (dereddener.hpp:41-62).  The three stretched
// medians are written straight into their zones of `median`, which equals
// the reference's copy sequence (the second copy's overrun past pos25 is
// overwritten by the third).
extern "C" int psx_deredden_median(const float* A, uint64_t nb, double bin_width, float b5, float b25,
                                   float* median, float* work, psx_stream s) {
  PSX_REQUIRE(A && median && work && nb >= 5 && bin_width > 0, "psx_deredden_median: bad arguments");
  const uint64_t c5 = nb / 5, c25 = nb / 5 / 5, c125 = nb / 5 / 5 / 5;
  PSX_REQUIRE(c125 >= 1, "psx_deredden_median: nb too small (%llu)", (unsigned long long)nb);
  float* m5 = work;
  float* m25 = m5 + c5 + 1;
  float* m125 = m25 + c25 + 1;
  int rc;
  if ((rc = psx_median_scrunch5(A, nb, m5, s))) return rc;
  if ((rc = psx_median_scrunch5(m5, c5, m25, s))) return rc;
  if ((rc = psx_median_scrunch5(m25, c25, m125, s))) return rc;
  const int64_t pos5 = (int64_t)(int)((double)b5 / bin_width);
  const int64_t pos25 = (int64_t)(int)((double)b25 / bin_width);
  auto clampz = [&](int64_t v) -> uint64_t { return v < 0 ? 0 : (v > (int64_t)nb ? nb : (uint64_t)v); };
  const uint64_t z0 = 0, z1 = clampz(pos5), z2 = clampz(pos25);
  const hipStream_t hs = as_stream(s);
  // zone [0, pos5) <- m5
  if (z1 > z0) {
    k_stretch_zone<<<psx::grid_for(z1 - z0, kBlock), kBlock, 0, hs>>>(m5, c5, median, z0, z1, stretch_step(c5, nb));
    PSX_CHECK_LAUNCH("k_stretch_zone");
  }
  // zone [pos5, pos25) <- m25 (the reference copies pos25 elements from pos5;
  // everything from pos25 on is rewritten by the third zone)
  const uint64_t z1e = clampz(pos5 + pos25);
  const uint64_t hi25 = z2 > z1 ? z2 : z1;
  if (z1e > z1) {
    const uint64_t hi = z1e < hi25 ? z1e : hi25;
    if (hi > z1) {
      k_stretch_zone<<<psx::grid_for(hi - z1, kBlock), kBlock, 0, hs>>>(m25, c25, median, z1, hi,
                                                                        stretch_step(c25, nb));
      PSX_CHECK_LAUNCH("k_stretch_zone");
    }
  }
  // zone [pos25, nb) <- m125
  if (nb > z2) {
    k_stretch_zone<<<psx::grid_for(nb - z2, kBlock), kBlock, 0, hs>>>(m125, c125, median, z2, nb,
                                                                      stretch_step(c125, nb));
    PSX_CHECK_LAUNCH("k_stretch_zone");
  }
  return psx::ok();
}

// ------------------------------------------------ running median, no allocation
// psx_deredden_median_noalloc: the same kernels, with the three scrunched
// medians kept in a file-scope device array instead of a caller workspace.
// The array's name is only a host-side handle: kernels get its device
// address, looked up once per device with hipGetSymbolAddress.  Every call on
// a device shares that device's array (calls that overlap on one device
// race); every element a call reads is written earlier in the same call, so
// nothing needs clearing.
namespace {
constexpr uint64_t kNoAllocMaxBins = (1ull << 25) + 1;  // 2^26-sample series
constexpr uint64_t kNoAllocScratchFloats = kNoAllocMaxBins / 5 + kNoAllocMaxBins / 25 + kNoAllocMaxBins / 125 + 3;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

static __device__ __align__(256) float scratch[kNoAllocScratchFloats];

static int scratch_on_current_device(float** out) {
  int dev = 0;
  PSX_CHECK_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return psx::fail(PSX_ERR_UNSUPPORTED, "device index %d out of range", dev);
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return psx::fail(PSX_ERR_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
    g_scratch_ptr[dev] = (float*)p;
  }
  *out = g_scratch_ptr[dev];
  return PSX_OK;
}

extern "C" uint64_t psx_deredden_noalloc_max_bins(void) { return kNoAllocMaxBins; }

extern "C" int psx_deredden_median_noalloc(const float* A, uint64_t nb, double bin_width, float b5, float b25,
                                           float* median, psx_stream s) {
  PSX_REQUIRE(nb <= kNoAllocMaxBins,
              "psx_deredden_median_noalloc: %llu bins exceed the static scratch (at most %llu bins)",
              (unsigned long long)nb, (unsigned long long)kNoAllocMaxBins);
  float* work = nullptr;
  const int rc = scratch_on_current_device(&work);
  if (rc) return rc;
  return psx_deredden_median(A, nb, bin_width, b5, b25, median, work, s);
}

extern "C" uint64_t psx_reduce_work_bytes(void) { return (kReduceBlocks + 1) * sizeof(double2); }

static int sum_sq(const float* x, uint64_t first, uint64_t n, void* work, psx_stream s, double* s1, double* s2) {
  double2* partial = (double2*)work;
  double2* out = partial + kReduceBlocks;
  const unsigned g = psx::grid_for(n - first, kBlock, kReduceBlocks);
  k_sum_sq<<<g, kBlock, 0, as_stream(s)>>>(x, first, n, partial, nullptr);
  PSX_CHECK_LAUNCH("k_sum_sq");
  k_sum_final<<<1, kBlock, 0, as_stream(s)>>>(partial, g, out, nullptr);
  PSX_CHECK_LAUNCH("k_sum_final");
  double2 h;
  PSX_CHECK_HIP(hipMemcpyAsync(&h, out, sizeof(h), hipMemcpyDeviceToHost, as_stream(s)));
  PSX_CHECK_HIP(hipStreamSynchronize(as_stream(s)));
  *s1 = h.x;
  *s2 = h.y;
  return PSX_OK;
}

// stats: GPU_rms then GPU_mean (src/kernels.cu:433-461): float(sum)/float(n).
extern "C" int psx_mean_rms(const float* x, uint64_t n, uint64_t first, float* mean, float* rms, void* work,
                            psx_stream s) {
  PSX_REQUIRE(x && work && n > first, "psx_mean_rms: bad arguments");
  double s1, s2;
  int rc = sum_sq(x, first, n, work, s, &s1, &s2);
  if (rc) return rc;
  const float cnt = (float)(int64_t)(n - first);
  if (mean) *mean = (float)s1 / cnt;
  if (rms) *rms = sqrtf((float)s2 / cnt);
  return psx::ok();
}

extern "C" int psx_normalise(float* p, float mean, float sigma, uint64_t n, psx_stream s) {
  PSX_REQUIRE(p, "psx_normalise: null pointer");
  PSX_REQUIRE(((uintptr_t)p & 15) == 0, "psx_normalise: buffer must be 16-byte aligned");
  if (!n) return psx::ok();
  k_normalise<<<psx::grid_for(n / 4 + 1, kBlock), kBlock, 0, as_stream(s)>>>(p, n, mean, sigma);
  PSX_CHECK_LAUNCH("k_normalise");
  return psx::ok();
}

namespace psx {
int sum_sq_async(const float* x, uint64_t first, uint64_t n, void* work, double2* out, hipStream_t s,
                 const BaselineState* skip) {
  double2* partial = (double2*)work;
  const unsigned g = psx::grid_for(n - first, kBlock, kReduceBlocks);
  k_sum_sq<<<g, kBlock, 0, s>>>(x, first, n, partial, skip);
  PSX_CHECK_LAUNCH("k_sum_sq");
  k_sum_final<<<1, kBlock, 0, s>>>(partial, g, out, skip);
  PSX_CHECK_LAUNCH("k_sum_final");
  return PSX_OK;
}

int remove_baseline_async(float* x, uint64_t n, void* work, BaselineState* st, int steps, hipStream_t s) {
  if (!x || !work || !st || n == 0 || steps < 1) return psx::fail(PSX_ERR_ARG, "remove_baseline_async: bad arguments");
  double2* sums = (double2*)work + kReduceBlocks;
  k_baseline_init<<<1, 64, 0, s>>>(st);
  PSX_CHECK_LAUNCH("k_baseline_init");
  for (int i = 0; i < steps; ++i) {
    int rc = sum_sq_async(x, 0, n, work, sums, s, st);
    if (rc) return rc;
    k_subtract_dev<<<psx::grid_for(n, kBlock), kBlock, 0, s>>>(x, n, sums, st);
    PSX_CHECK_LAUNCH("k_subtract_dev");
    k_baseline_step<<<1, 64, 0, s>>>(n, sums, st);
    PSX_CHECK_LAUNCH("k_baseline_step");
  }
  return PSX_OK;
}
}  // namespace psx

// GPU_remove_baseline (src/kernels.cu:463-479).
extern "C" int psx_remove_baseline(float* x, uint64_t n, void* work, psx_stream s) {
  PSX_REQUIRE(x && work && n > 0, "psx_remove_baseline: bad arguments");
  int count = 0;
  float mean;
  do {
    double s1, s2;
    int rc = sum_sq(x, 0, n, work, s, &s1, &s2);
    if (rc) return rc;
    mean = (float)s1 / (float)(int64_t)n;
    k_subtract<<<psx::grid_for(n, kBlock), kBlock, 0, as_stream(s)>>>(x, n, mean);
    PSX_CHECK_LAUNCH("k_subtract");
    if (count++ > 100) break;
  } while (fabsf(mean) > 5e-7 * (double)n);
  PSX_CHECK_HIP(hipStreamSynchronize(as_stream(s)));
  return psx::ok();
}

extern "C" int psx_fill(float* x, uint64_t start, uint64_t end, float value, psx_stream s) {
  PSX_REQUIRE(x, "psx_fill: null pointer");
  if (end <= start) return psx::ok();
  k_fill<<<psx::grid_for(end - start, kBlock), kBlock, 0, as_stream(s)>>>(x, start, end, value);
  PSX_CHECK_LAUNCH("k_fill");
  return psx::ok();
}

extern "C" int psx_convert_to_f32(const void* in, int in_type, float* out, uint64_t n, psx_stream s) {
  PSX_REQUIRE(in && out, "psx_convert_to_f32: null pointer");
  if (!n) return psx::ok();
  const unsigned g = psx::grid_for(n, kBlock);
  const hipStream_t hs = as_stream(s);
  switch (in_type) {
    case PSX_U8: k_convert<uint8_t><<<g, kBlock, 0, hs>>>((const uint8_t*)in, out, n); break;
    case PSX_I8: k_convert<int8_t><<<g, kBlock, 0, hs>>>((const int8_t*)in, out, n); break;
    case PSX_U32: k_convert<uint32_t><<<g, kBlock, 0, hs>>>((const uint32_t*)in, out, n); break;
    case PSX_F32: k_convert<float><<<g, kBlock, 0, hs>>>((const float*)in, out, n); break;
    default: return psx::fail(PSX_ERR_ARG, "psx_convert_to_f32: unknown input type %d", in_type);
  }
  PSX_CHECK_LAUNCH("k_convert");
  return psx::ok();
}

extern "C" int psx_harmonic_sum(const float* P, float* const* folds, uint32_t nfolds, uint64_t nb, psx_stream s) {
  PSX_REQUIRE(P && folds, "psx_harmonic_sum: null pointer");
  PSX_REQUIRE(nfolds >= 1 && nfolds <= 5, "psx_harmonic_sum: nfolds must be 1..5 (got %u)", nfolds);
  FoldPtrs fp = {};
  for (uint32_t k = 0; k < nfolds; ++k) {
    PSX_REQUIRE(folds[k], "psx_harmonic_sum: null fold pointer %u", k);
    fp.f[k] = folds[k];
  }
  if (!nb) return psx::ok();
  psx::TraceRange range("Harmonic summing");
  const unsigned g = psx::grid_for(nb, kBlock);
  const hipStream_t hs = as_stream(s);
  switch (nfolds) {
    case 1: k_harmonic<1><<<g, kBlock, 0, hs>>>(P, fp, nb); break;
    case 2: k_harmonic<2><<<g, kBlock, 0, hs>>>(P, fp, nb); break;
    case 3: k_harmonic<3><<<g, kBlock, 0, hs>>>(P, fp, nb); break;
    case 4: k_harmonic<4><<<g, kBlock, 0, hs>>>(P, fp, nb); break;
    default: k_harmonic<5><<<g, kBlock, 0, hs>>>(P, fp, nb); break;
  }
  PSX_CHECK_LAUNCH("k_harmonic");
  return psx::ok();
}

extern "C" uint64_t psx_find_peaks_work_bytes(uint64_t n) {
  return ((n + kPeakTile - 1) / kPeakTile + 1) * sizeof(uint32_t);
}

extern "C" int psx_find_peaks(const float* p, int64_t start, int64_t end, float thr, int32_t* d_idx,
                              float* d_snr, uint32_t cap, uint32_t* d_count, void* work, psx_stream s) {
  PSX_REQUIRE(p && d_count && work, "psx_find_peaks: null pointer");
  PSX_REQUIRE(cap == 0 || (d_idx && d_snr), "psx_find_peaks: null output arrays");
  if (start < 0) start = 0;
  const hipStream_t hs = as_stream(s);
  if (end <= start) {
    k_set_u32<<<1, 64, 0, hs>>>(d_count, 0u);
    PSX_CHECK_LAUNCH("k_set_u32");
    return psx::ok();
  }
  const uint64_t n = (uint64_t)(end - start);
  PSX_REQUIRE(n < (1ull << 31), "psx_find_peaks: range exceeds 2^31 bins");
  const unsigned ntiles = (unsigned)((n + kPeakTile - 1) / kPeakTile);
  uint32_t* counts = (uint32_t*)work;
  k_peak_count<<<ntiles, kBlock, 0, hs>>>(p, start, end, thr, counts);
  PSX_CHECK_LAUNCH("k_peak_count");
  k_peak_write<<<ntiles, kBlock, 0, hs>>>(p, start, end, thr, counts, ntiles, d_idx, d_snr, cap, d_count);
  PSX_CHECK_LAUNCH("k_peak_write");
  return psx::ok();
}
