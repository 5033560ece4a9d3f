// fft_dev.hpp - device building blocks of the FP32 Stockham FFT (gfx950).
//
// A "pass" transforms C columns of length L per workgroup.  Column c,
// element e lives at  base + c*col_stride + e*elem_stride  in global memory;
// the transform runs as radix-R Stockham autosort stages (R in {16,8,4,2,3})
// whose intermediate results go through LDS.  Each thread owns P points in
// registers; stage s assigns butterfly slot b of thread t to
// g = b*NT + t, mapped either column-fast (c = g % C) for the global side
// whose columns are adjacent in memory, or element-fast (j = g % (L/R)).
// Only forward transforms are instantiated: the inverse is computed as
// conj(FFT(conj(x))), with the conjugations folded into the neighbouring
// passes.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

namespace psx {
namespace fft {

// ------------------------------------------------------------ radix plans
__host__ __device__ constexpr int ilog2c(int v) { return v <= 1 ? 0 : 1 + ilog2c(v / 2); }
__host__ __device__ constexpr bool has3(int L) { return L % 3 == 0; }
__host__ __device__ constexpr int pow2part(int L) { return has3(L) ? L / 3 : L; }
// Radix list: [3?] then as many 16s as fit, then the 2^r remainder.
__host__ __device__ constexpr int num_stages(int L) {
  return (has3(L) ? 1 : 0) + ilog2c(pow2part(L)) / 4 + ((ilog2c(pow2part(L)) % 4) ? 1 : 0);
}
__host__ __device__ constexpr int radix_at(int L, int s) {
  return (has3(L) && s == 0) ? 3
         : ((s - (has3(L) ? 1 : 0)) < ilog2c(pow2part(L)) / 4) ? 16
                                                                : (1 << (ilog2c(pow2part(L)) % 4));
}
__host__ __device__ constexpr int max_pow2_radix(int L) {
  return ilog2c(pow2part(L)) >= 4 ? 16 : (1 << ilog2c(pow2part(L)));
}
// Points per thread: a multiple of every radix of the plan.
__host__ __device__ constexpr int points_per_thread(int L) {
  return has3(L) ? 3 * max_pow2_radix(L) : max_pow2_radix(L);
}
__host__ __device__ constexpr int ns_before(int L, int s) { return s == 0 ? 1 : ns_before(L, s - 1) * radix_at(L, s - 1); }
// LDS row pitch in float2: one pad slot per 16 elements plus one per row.
__host__ __device__ constexpr int lds_pitch(int L, int C) { return L + L / 16 + (C > 1 ? 1 : 0); }
// Passes over strided columns exchange one float plane at a time between
// stages (real parts, then imaginary parts): one float of LDS per point,
// half of a complex exchange, so twice the workgroups fit on a CU.  Row
// passes exchange whole complex values (measured faster there: their extra
// barriers cost more than the occupancy gains).
__host__ __device__ constexpr bool split_planes(int mode) { return mode == 1; }
__host__ __device__ constexpr int lds_bytes(int L, int C, bool split) {
  return C * lds_pitch(L, C) * (split ? 4 : 8);
}
// Waves per SIMD that the LDS footprint allows for a pass (160 KB of LDS and
// at most 32 waves per CU, 4 SIMDs): the register budget the kernel is
// compiled for (512 / waves VGPRs), so that registers never cap the
// workgroups per CU below what the LDS admits.  One resident workgroup per
// CU idles the CU at every barrier and every load phase; two overlap them.
// Plans with a radix-3 stage hold 48 points per thread (96 VGPRs of data)
// and keep the compiler's own register allocation (target 1).
__host__ __device__ constexpr int wgs_per_cu(int L, int C, int bytes) {
  const int nt = C * L / points_per_thread(L);
  const int by_lds = 163840 / bytes;
  return by_lds * nt > 2048 ? 2048 / nt : by_lds;
}
__host__ __device__ constexpr int waves_per_simd(int L, int C, int mode) {
  const int nt = C * L / points_per_thread(L);
  const int w = (wgs_per_cu(L, C, lds_bytes(L, C, split_planes(mode))) * ((nt + 63) / 64) + 3) / 4;
  return points_per_thread(L) > 16 ? 1 : (w > 8 ? 8 : (w < 1 ? 1 : w));
}
// Where a pass reads its stage twiddles W_L^e from.  Loading them from the
// global table at the butterfly that uses them stalls on the load latency
// after every exchange, so:
//   TW_LDS       the L-entry table is copied into LDS behind the exchange
//                buffer when that leaves the workgroups per CU unchanged;
//   TW_LDS_HALF  else its first L/2 entries, when those fit that way
//                (W_L^(m + L/2) = -W_L^m: one sign flip per lookup);
//   TW_PREFETCH  otherwise, when the register budget is at least 128 VGPRs
//                and a thread holds at most 16 points, each stage's twiddles
//                are loaded before the preceding LDS exchange (their latency
//                hides behind it);
//   TW_GLOBAL    loads at the point of use.
enum TwiddleSource { TW_GLOBAL = 0, TW_LDS = 1, TW_PREFETCH = 2, TW_LDS_HALF = 3 };
// LDS twiddle tables keep entry m at slot m + m/32: a stage reads W[k*r*TS]
// for consecutive k, i.e. strides of r*TS entries, which without the pad put
// a wave's lanes on a few banks (all of them on one bank for 16-point
// spans); with it the reads average under the 4 cycles of a conflict-free
// 64-bit access in a bank model of the access patterns.
__host__ __device__ constexpr int tw_slot(int m) { return m + (m >> 5); }
__host__ __device__ constexpr int tw_table_bytes(int entries) { return 8 * tw_slot(entries); }
__host__ __device__ constexpr int twiddle_source(int L, int C, int mode) {
  const int base = lds_bytes(L, C, split_planes(mode));
  if (num_stages(L) < 2) return TW_GLOBAL;
  const int full = tw_table_bytes(L), half = tw_table_bytes(L / 2);
  if (base + full <= 163840 && wgs_per_cu(L, C, base + full) == wgs_per_cu(L, C, base)) return TW_LDS;
  if (base + half <= 163840 && wgs_per_cu(L, C, base + half) == wgs_per_cu(L, C, base)) return TW_LDS_HALF;
  if (points_per_thread(L) <= 16 && 512 / waves_per_simd(L, C, mode) >= 128) return TW_PREFETCH;
  return TW_GLOBAL;
}
// Dynamic LDS of a pass kernel: the exchange buffer plus the twiddle table.
__host__ __device__ constexpr int pass_lds_bytes(int L, int C, int mode) {
  return lds_bytes(L, C, split_planes(mode)) +
         (twiddle_source(L, C, mode) == TW_LDS
              ? tw_table_bytes(L)
              : (twiddle_source(L, C, mode) == TW_LDS_HALF ? tw_table_bytes(L / 2) : 0));
}

// ------------------------------------------------- global memory access
// Loads and stores at a workgroup-uniform base plus a 32-bit element index.
// The base is cast to the global address space and the byte offset is a
// zero-extended 32-bit value, which selects the SGPR-base + 32-bit
// VGPR-offset form of global_load/global_store: one VGPR per address and no
// 64-bit address arithmetic (generic pointers otherwise become flat
// accesses with 64-bit VGPR addresses).  Offsets stay below 2^32 bytes: an
// index is below n <= 2^28 samples.
#define PSX_GLOBAL __attribute__((address_space(1)))
typedef float psx_f2v __attribute__((ext_vector_type(2)));

__device__ inline float2 ldg2(const float2* base, uint32_t i) {
  const psx_f2v v = *(const PSX_GLOBAL psx_f2v*)((const PSX_GLOBAL char*)base + (uint64_t)(i * 8u));
  return make_float2(v.x, v.y);
}
__device__ inline float ldg1(const float* base, uint32_t i) {
  return *(const PSX_GLOBAL float*)((const PSX_GLOBAL char*)base + (uint64_t)(i * 4u));
}
// floats i and i+1 with one 8-byte load that needs only 4-byte alignment
typedef float psx_f2u __attribute__((ext_vector_type(2), aligned(4)));
__device__ inline float2 ldg2u(const float* base, uint32_t i) {
  const psx_f2u v = *(const PSX_GLOBAL psx_f2u*)((const PSX_GLOBAL char*)base + (uint64_t)(i * 4u));
  return make_float2(v.x, v.y);
}
__device__ inline void stg1(float* base, uint32_t i, float v) {
  *(PSX_GLOBAL float*)((PSX_GLOBAL char*)base + (uint64_t)(i * 4u)) = v;
}
__device__ inline void stg2(float2* base, uint32_t i, float2 v) {
  psx_f2v w;
  w.x = v.x;
  w.y = v.y;
  *(PSX_GLOBAL psx_f2v*)((PSX_GLOBAL char*)base + (uint64_t)(i * 8u)) = w;
}
// Streaming (non-temporal) load: the pass reads each element once.
__device__ inline float2 ldg2_nt(const float2* base, uint32_t i) {
  const psx_f2v v = __builtin_nontemporal_load((const PSX_GLOBAL psx_f2v*)((const PSX_GLOBAL char*)base +
                                                                            (uint64_t)(i * 8u)));
  return make_float2(v.x, v.y);
}
// Streaming (non-temporal) store: the line is not kept dirty in the caches
// for a later write-back that would compete with the next kernel's traffic.
// Used where whole 128-byte lines are written (column tiles of 16).
__device__ inline void stg2_nt(float2* base, uint32_t i, float2 v) {
  psx_f2v w;
  w.x = v.x;
  w.y = v.y;
  __builtin_nontemporal_store(w, (PSX_GLOBAL psx_f2v*)((PSX_GLOBAL char*)base + (uint64_t)(i * 8u)));
}

// ---------------------------------------------------------- complex math
__device__ inline float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ inline float2 cmul(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x));
}
// a * (-i)
__device__ inline float2 mul_mi(float2 a) { return make_float2(a.y, -a.x); }
__device__ inline float2 conjf2(float2 a) { return make_float2(a.x, -a.y); }

// Forward DFTs (exp(-2 pi i rk/R)), in place, natural order in and out.
__device__ inline void dft2(float2& a, float2& b) {
  const float2 t = a;
  a = cadd(t, b);
  b = csub(t, b);
}

__device__ inline void dft4(float2& v0, float2& v1, float2& v2, float2& v3) {
  const float2 a0 = cadd(v0, v2), a1 = csub(v0, v2), a2 = cadd(v1, v3), a3 = mul_mi(csub(v1, v3));
  v0 = cadd(a0, a2);
  v2 = csub(a0, a2);
  v1 = cadd(a1, a3);
  v3 = csub(a1, a3);
}

__device__ inline void dft3(float2& v0, float2& v1, float2& v2) {
  const float s3 = 0.866025403784438646763723170752936f;
  const float2 t1 = cadd(v1, v2), t2 = csub(v1, v2);
  const float2 base = make_float2(v0.x - 0.5f * t1.x, v0.y - 0.5f * t1.y);
  // -i * s3 * t2
  const float2 rot = make_float2(s3 * t2.y, -s3 * t2.x);
  v0 = cadd(v0, t1);
  v1 = cadd(base, rot);
  v2 = csub(base, rot);
}

// W_16^m = exp(-2 pi i m/16), m = 0..9 (the products q*k1 of the 4x4 split).
__device__ inline float2 w16(int m) {
  const float c1 = 0.923879532511286756128183189396788f;  // cos(pi/8)
  const float s1 = 0.382683432365089771728459984030399f;  // sin(pi/8)
  const float h = 0.707106781186547524400844362104849f;
  switch (m) {
    case 0: return make_float2(1.f, 0.f);
    case 1: return make_float2(c1, -s1);
    case 2: return make_float2(h, -h);
    case 3: return make_float2(s1, -c1);
    case 4: return make_float2(0.f, -1.f);
    case 5: return make_float2(-s1, -c1);
    case 6: return make_float2(-h, -h);
    case 7: return make_float2(-c1, -s1);
    case 8: return make_float2(-1.f, 0.f);
    default: return make_float2(-c1, s1);  // 9
  }
}

// DFT8 = 2 x 4: n = q + 2r, k = k1 + 4 k2.
__device__ inline void dft8(float2* v) {
  float2 a[2][4];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    a[q][0] = v[q];
    a[q][1] = v[q + 2];
    a[q][2] = v[q + 4];
    a[q][3] = v[q + 6];
    dft4(a[q][0], a[q][1], a[q][2], a[q][3]);
  }
  const float h = 0.707106781186547524400844362104849f;
  a[1][1] = cmul(a[1][1], make_float2(h, -h));
  a[1][2] = mul_mi(a[1][2]);
  a[1][3] = cmul(a[1][3], make_float2(-h, -h));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 x0 = a[0][k1], x1 = a[1][k1];
    dft2(x0, x1);
    v[k1] = x0;
    v[k1 + 4] = x1;
  }
}

// DFT16 = 4 x 4: n = q + 4r, k = k1 + 4 k2, twiddle W16^(q*k1).
__device__ inline void dft16(float2* v) {
  float2 a[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    a[q][0] = v[q];
    a[q][1] = v[q + 4];
    a[q][2] = v[q + 8];
    a[q][3] = v[q + 12];
    dft4(a[q][0], a[q][1], a[q][2], a[q][3]);
  }
#pragma unroll
  for (int q = 1; q < 4; ++q)
#pragma unroll
    for (int k1 = 1; k1 < 4; ++k1) a[q][k1] = (q * k1 == 4) ? mul_mi(a[q][k1]) : cmul(a[q][k1], w16(q * k1));
#pragma unroll
  for (int k1 = 0; k1 < 4; ++k1) {
    float2 x0 = a[0][k1], x1 = a[1][k1], x2 = a[2][k1], x3 = a[3][k1];
    dft4(x0, x1, x2, x3);
    v[k1] = x0;
    v[k1 + 4] = x1;
    v[k1 + 8] = x2;
    v[k1 + 12] = x3;
  }
}

template <int R>
__device__ inline void dft(float2* v) {
  if constexpr (R == 2) dft2(v[0], v[1]);
  else if constexpr (R == 3) dft3(v[0], v[1], v[2]);
  else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
  else if constexpr (R == 8) dft8(v);
  else dft16(v);
}

// ------------------------------------------------------------ pass args
// One acceleration trial of a batch: the dereddened series of its DM and the
// resampling factor af (device_resampleII, src/kernels.cu:361), plus the
// normalisation constants mean*N, std*N (pipeline_multi.cu:271).
struct TrialDesc {
  const float* tim;
  double af;
  float mean_n;
  float std_n;
  uint32_t slot;
  uint32_t pad;
};

struct PassArgs {
  const float2* in;
  float2* out;
  int64_t in_col, in_elem, in_sub, in_tr;
  int64_t out_col, out_elem, out_sub, out_tr;
  const float2* twL;     // W_L^e, e in [0, L)
  const float2* tw4_lo;  // W_Mtw^(x & mask)
  const float2* tw4_hi;  // W_Mtw^((x >> shift) << shift)
  uint32_t tw4_shift;
  uint32_t tw4_mask;
  const float2* tw4_r;  // W_(Mtw*R/L)^k, R the last stage's radix
  int32_t tw4;       // multiply output (c, e) by W_Mtw^(c*e)
  int32_t in_cf;     // stage-0 mapping column-fast
  int32_t out_cf;    // last-stage mapping column-fast
  int32_t in_conj;   // conjugate inputs on load
  int32_t out_conj;  // conjugate outputs on store
  // Fused resampling (first pass of a batched search): complex input m of
  // trial blockIdx.z is (tim[idx(2m)], tim[idx(2m+1)]) with the resample index
  // of the trial's af over a series of rs_n samples.
  int32_t resample;
  int32_t rs_tile_major;  // block order of the resampling pass (see k_fft_pass)
  uint64_t rs_n;
  const TrialDesc* descs;
  // Several column tiles per workgroup (kernels where two_tiles_ok): the
  // grid covers ceil(ntiles/tiles) tiles and workgroup b also does tiles
  // b + gridDim.x, b + 2*gridDim.x, ... (< ntiles), each tile's loads issued
  // before the previous one is transformed.  0 or 1: one tile.
  int32_t tiles;
  uint32_t ntiles;
};
// Passes that may take two tiles per workgroup: contiguous tiles of at
// least 4 rows whose register budget (at least 128 VGPRs) holds a second
// tile's points while the first is transformed (narrower tiles spill).
__host__ __device__ constexpr bool two_tiles_ok(int L, int C, int mode) {
  return mode == 0 && C >= 4 && points_per_thread(L) <= 16 && waves_per_simd(L, C, mode) <= 4;
}

__device__ inline float2 tw4_lookup(const PassArgs& a, uint64_t x) {
  const float2 lo = a.tw4_lo[x & a.tw4_mask];
  const float2 hi = a.tw4_hi[x >> a.tw4_shift];
  return cmul(hi, lo);
}

template <int L, int C>
__device__ inline int lds_index(int c, int e) {
  return c * lds_pitch(L, C) + e + (e >> 4);
}
// Index step between the elements e0 + r*K (r < N) of one butterfly's run,
// or 0 when it is not constant.  Runs are aligned to K*N (Stockham
// indexing), so a run moves by whole padding groups when K is a multiple of
// 16 (step K + K/16) and stays inside one 16-element group when K*N divides
// 16 (step K).  A constant step becomes an instruction offset on one base
// address instead of a separate address per element.
__host__ __device__ constexpr int lds_step(int K, int N) {
  return (K % 16 == 0) ? K + K / 16 : ((K * N <= 16 && 16 % (K * N) == 0) ? K : 0);
}

template <int L, int C, int R, int NT>
__device__ inline void bf_coords(int b, int t, bool cf, int& c, int& j) {
  const int g = b * NT + t;
  constexpr int NBC = L / R;
  if (C == 1) {
    c = 0;
    j = g;
  } else if (cf) {
    c = g % C;
    j = g / C;
  } else {
    c = g / NBC;
    j = g % NBC;
  }
}

}  // namespace fft
}  // namespace psx
