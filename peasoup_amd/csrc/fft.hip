// fft.hip - batched FP32 R2C / C2R FFT for gfx950 (replaces cuFFT behind
// CuFFTerR2C / CuFFTerC2R, include/transforms/ffter.hpp:54-110).
//
// Real length n = 2M is computed as an M-point complex FFT of
// z[m] = x[2m] + i x[2m+1] followed by an in-place "untangle" pass that
// yields the n/2+1 non-redundant bins; C2R runs the inverse steps.  The
// complex FFT is a 1-, 2- or 3-pass four-step decomposition (M = M1*M2 or
// Q*M2*M1); every pass is one launch of k_fft_pass, which reads its columns
// once, runs all Stockham stages through LDS, applies the inter-pass twiddle
// and writes once, so a pass costs one read and one write of the array.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "fft_dev.hpp"
#include "fft_plan.hpp"
#include "kernels_dev.hpp"

using psx::as_stream;
using namespace psx::fft;

namespace {

// ------------------------------------------------------------ pass kernel
// Pass modes fix, at compile time, which side of the pass has adjacent
// columns (column-fast thread mapping, C*8 contiguous bytes per row) and
// whether the four-step twiddle is applied on the way out:
//   PM_CONTIG        contiguous rows in and out (single-pass transforms)
//   PM_COLS_TW       strided columns in and out, twiddle W_M^(c*e) applied
//   PM_ROWS_TO_COLS  contiguous rows in, strided columns out
// RS: the first read goes through the resampler (batched search).
enum PassMode { PM_CONTIG = 0, PM_COLS_TW = 1, PM_ROWS_TO_COLS = 2 };

template <int MODE>
struct ModeTraits {
  static constexpr bool in_cf = MODE == PM_COLS_TW;
  static constexpr bool out_cf = MODE != PM_CONTIG;
  static constexpr bool tw4 = MODE == PM_COLS_TW;
};

template <int S, int L, int C, int MODE>
__device__ inline void run_stage(float2 (&v)[points_per_thread(L)], float2 (&tw)[points_per_thread(L)], float* lds,
                                 const PassArgs& a, float2* out, unsigned col0, int t) {
  constexpr int P = points_per_thread(L);
  constexpr int NT = C * L / P;
  constexpr int NSTG = num_stages(L);
  constexpr int R = radix_at(L, S);
  constexpr int NB = P / R;
  constexpr int Ns = ns_before(L, S);
  constexpr bool LAST = (S == NSTG - 1);
  constexpr bool cf = (S == 0) ? ModeTraits<MODE>::in_cf : (LAST ? ModeTraits<MODE>::out_cf : false);
  constexpr int TWS = twiddle_source(L, C, MODE);

  // twiddle + butterfly
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    int c, j;
    bf_coords<L, C, R, NT>(b, t, cf, c, j);
    if (S > 0) {
      const int k = j % Ns;
      constexpr int TS = L / (Ns * R);
#pragma unroll
      for (int r = 1; r < R; ++r) {
        float2 w;
        if constexpr (TWS == TW_PREFETCH)
          w = tw[b * (R - 1) + r - 1];
        else if constexpr (TWS == TW_LDS)
          w = reinterpret_cast<const float2*>(lds + lds_bytes(L, C, split_planes(MODE)) / 4)[tw_slot(k * r * TS)];
        else if constexpr (TWS == TW_LDS_HALF) {
          const int m = k * r * TS;
          const bool upper = m >= L / 2;
          w = reinterpret_cast<const float2*>(lds + lds_bytes(L, C, split_planes(MODE)) / 4)[tw_slot(upper ? m - L / 2
                                                                                                         : m)];
          if (upper) w = make_float2(-w.x, -w.y);
        }
        else
          w = ldg2(a.twL, (uint32_t)(k * r * TS));
        v[b * R + r] = cmul(v[b * R + r], w);
      }
    }
    dft<R>(&v[b * R]);
  }

  if constexpr (!LAST) {
    constexpr int R2 = radix_at(L, S + 1);
    constexpr int NB2 = P / R2;
    constexpr bool cf2 = (S + 1 == NSTG - 1) ? ModeTraits<MODE>::out_cf : false;
    if constexpr (TWS == TW_PREFETCH) {
      // next stage's twiddles, in flight during the exchange below
      constexpr int Ns2 = ns_before(L, S + 1);
      constexpr int TS2 = L / (Ns2 * R2);
#pragma unroll
      for (int b = 0; b < NB2; ++b) {
        int c, j;
        bf_coords<L, C, R2, NT>(b, t, cf2, c, j);
        const int k = j % Ns2;
#pragma unroll
        for (int r = 1; r < R2; ++r) tw[b * (R2 - 1) + r - 1] = ldg2(a.twL, (uint32_t)(k * r * TS2));
      }
    }
    constexpr int WS = lds_step(Ns, R);            // write run: idxD + r*Ns
    constexpr int RSTEP = lds_step(L / R2, R2);    // read run: j + r*(L/R2)
    if constexpr (!split_planes(MODE)) {
      float2* lds2 = reinterpret_cast<float2*>(lds);
      if (S > 0) __syncthreads();
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        int c, j;
        bf_coords<L, C, R, NT>(b, t, cf, c, j);
        const int k = j % Ns;
        const int idxD = (j / Ns) * Ns * R + k;
        const int w0 = lds_index<L, C>(c, idxD);
#pragma unroll
        for (int r = 0; r < R; ++r) lds2[WS ? w0 + r * WS : lds_index<L, C>(c, idxD + r * Ns)] = v[b * R + r];
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NB2; ++b) {
        int c, j;
        bf_coords<L, C, R2, NT>(b, t, cf2, c, j);
        const int r0 = lds_index<L, C>(c, j);
#pragma unroll
        for (int r = 0; r < R2; ++r) v[b * R2 + r] = lds2[RSTEP ? r0 + r * RSTEP : lds_index<L, C>(c, j + r * (L / R2))];
      }
    } else {
    // real plane, then imaginary plane: each point's component is written
    // before it is overwritten by the read of the same component
#pragma unroll
    for (int plane = 0; plane < 2; ++plane) {
      if (S > 0 || plane > 0) __syncthreads();
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        int c, j;
        bf_coords<L, C, R, NT>(b, t, cf, c, j);
        const int k = j % Ns;
        const int idxD = (j / Ns) * Ns * R + k;
        const int w0 = lds_index<L, C>(c, idxD);
#pragma unroll
        for (int r = 0; r < R; ++r)
          lds[WS ? w0 + r * WS : lds_index<L, C>(c, idxD + r * Ns)] = plane ? v[b * R + r].y : v[b * R + r].x;
      }
      __syncthreads();
#pragma unroll
      for (int b = 0; b < NB2; ++b) {
        int c, j;
        bf_coords<L, C, R2, NT>(b, t, cf2, c, j);
        const int r0 = lds_index<L, C>(c, j);
#pragma unroll
        for (int r = 0; r < R2; ++r) {
          const float x = lds[RSTEP ? r0 + r * RSTEP : lds_index<L, C>(c, j + r * (L / R2))];
          if (plane)
            v[b * R2 + r].y = x;
          else
            v[b * R2 + r].x = x;
        }
      }
    }
    }
    run_stage<S + 1, L, C, MODE>(v, tw, lds, a, out, col0, t);
  } else {
    // element offsets inside one transform are < n/2 <= 2^27: 32-bit math
    const uint32_t ocol = (uint32_t)a.out_col, oel = (uint32_t)a.out_elem;
    // four-step twiddle W_Mtw^(cg*e) with e = j + r*(L/R): for a last radix
    // R <= 8, the factor W_Mtw^(cg*j) from the two-level table once per
    // butterfly, times W_Mtw^(cg*r*L/R) = W_(Mtw*R/L)^(cg*r), which depends
    // on the column only (the column-fast mapping keeps a thread's column
    // for every b): R - 1 loads per thread instead of two per point.  A
    // radix-16 last stage would hold 16 such factors and spill at the
    // 64-VGPR budget: it looks up each point's twiddle whole.
    constexpr bool kSplitTw = ModeTraits<MODE>::tw4 && R <= 8;
    float2 wr[R];
    if constexpr (kSplitTw) {
      static_assert(cf, "the four-step twiddle needs the column-fast output mapping");
      int c0, j0;
      bf_coords<L, C, R, NT>(0, t, cf, c0, j0);
      const uint32_t cg0 = col0 + (uint32_t)c0;
      wr[0] = make_float2(1.f, 0.f);
#pragma unroll
      for (int r = 1; r < R; ++r) wr[r] = ldg2(a.tw4_r, cg0 * (uint32_t)r);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      int c, j;
      bf_coords<L, C, R, NT>(b, t, cf, c, j);
      const uint32_t cg = col0 + (uint32_t)c;
      const uint32_t o0 = cg * ocol + (uint32_t)j * oel;
      float2 wj = make_float2(1.f, 0.f);
      if constexpr (kSplitTw) {
        const uint32_t x = cg * (uint32_t)j;
        wj = cmul(ldg2(a.tw4_hi, x >> a.tw4_shift), ldg2(a.tw4_lo, x & a.tw4_mask));
      }
#pragma unroll
      for (int r = 0; r < R; ++r) {
        float2 val = v[b * R + r];
        if constexpr (kSplitTw) {
          val = cmul(val, r == 0 ? wj : cmul(wj, wr[r]));
        } else if constexpr (ModeTraits<MODE>::tw4) {
          const uint32_t x = cg * (uint32_t)(j + r * (L / R));
          val = cmul(val, cmul(ldg2(a.tw4_hi, x >> a.tw4_shift), ldg2(a.tw4_lo, x & a.tw4_mask)));
        }
        if (a.out_conj) val.y = -val.y;
        // column passes write whole lines (C*8 bytes per row, 128 for C=16)
        // and contiguous-row passes whole rows, each read once by the next
        // pass: streaming stores (measured 3% faster end to end at 2^23 for
        // the column passes and 1.5% more for the row passes; the 32-byte
        // runs of the rows-to-columns pass are 2x slower that way)
        if constexpr (MODE == PM_COLS_TW || MODE == PM_CONTIG)
          stg2_nt(out, o0 + (uint32_t)(r * (L / R)) * oel, val);
        else
          stg2(out, o0 + (uint32_t)(r * (L / R)) * oel, val);
      }
    }
  }
}

template <int L, int C, int MODE, bool RS>
__global__ void __launch_bounds__(C * L / points_per_thread(L))
__attribute__((amdgpu_waves_per_eu(waves_per_simd(L, C, MODE)))) k_fft_pass(const PassArgs a) {
  constexpr int P = points_per_thread(L);
  constexpr int NT = C * L / P;
  constexpr int R0 = radix_at(L, 0);
  constexpr int NB0 = P / R0;
  constexpr bool cf = ModeTraits<MODE>::in_cf;
  extern __shared__ float lds[];
  // a.rs_tile_major (resampling first pass, gridDim.y == 1): the trials of
  // one tile on consecutive hardware blocks, so that the trials of a batch
  // resample the same stretch of their DM's series at about the same time
  // and the re-reads hit the caches (trial-major order reads a series too
  // large for them, such as a 2^26-sample one, from HBM again for every
  // trial; for smaller series trial-major order is the faster one).
  unsigned blk, tr;
  if (RS && a.rs_tile_major) {
    const unsigned hw = blockIdx.x + gridDim.x * blockIdx.z;
    tr = hw % gridDim.z;
    blk = hw / gridDim.z;
  } else {
    blk = psx::xcd_logical_block(blockIdx.x, gridDim.x);
    tr = blockIdx.z;
  }
  const unsigned col0 = blk * C;
  float2* out = a.out + (int64_t)blockIdx.y * a.out_sub + (int64_t)tr * a.out_tr;
  const int t = threadIdx.x;
  const uint32_t icol = (uint32_t)a.in_col, iel = (uint32_t)a.in_elem;
  float2 v[P];
  float2 tw[P];
  if constexpr (twiddle_source(L, C, MODE) == TW_LDS || twiddle_source(L, C, MODE) == TW_LDS_HALF) {
    // the first use follows the stage-0 exchange and its barriers
    constexpr int NTW = twiddle_source(L, C, MODE) == TW_LDS ? L : L / 2;
    float2* ltw = reinterpret_cast<float2*>(lds + lds_bytes(L, C, split_planes(MODE)) / 4);
    for (int i = t; i < NTW; i += NT) ltw[tw_slot(i)] = ldg2(a.twL, (uint32_t)i);
  }
  if constexpr (RS) {
    // resampleII fused into the first read (src/kernels.cu:342-353); the
    // complex input m is the sample pair (2m, 2m+1) of the resampled series
    const TrialDesc d = a.descs[tr];
    const float* __restrict__ tim = d.tim;
    const double af = d.af;
    const uint32_t n = (uint32_t)a.rs_n;
    const double dn = (double)n;
    const uint32_t sub = (uint32_t)((int64_t)blockIdx.y * a.in_sub);
#pragma unroll
    for (int b = 0; b < NB0; ++b) {
      int c, j;
      bf_coords<L, C, R0, NT>(b, t, cf, c, j);
      const uint32_t m0 = sub + (col0 + (uint32_t)c) * icol + (uint32_t)j * iel;
#pragma unroll
      for (int r = 0; r < R0; ++r) {
        const uint32_t m = m0 + (uint32_t)(r * (L / R0)) * iel;
        // the pair's source samples are almost always adjacent (the index
        // slips by one only every 1/|af (2m - n)| samples): one 8-byte load
        // then, two 4-byte loads at a slip (3.8% faster pass than two loads
        // everywhere; both indices are in [0, n))
        const uint32_t j0 = psx::resample_index32(2 * m, n, dn, af);
        const uint32_t j1 = psx::resample_index32(2 * m + 1, n, dn, af);
        if (j1 == j0 + 1)
          v[b * R0 + r] = ldg2u(tim, j0);
        else
          v[b * R0 + r] = make_float2(ldg1(tim, j0), ldg1(tim, j1));
      }
    }
  } else {
    // every input element is read once: streaming loads (0.4% end to end
    // at 2^23 on top of the streaming row stores)
    const float2* in = a.in + (int64_t)blockIdx.y * a.in_sub + (int64_t)tr * a.in_tr;
    auto load = [&](unsigned cbase, float2(&dst)[P]) {
#pragma unroll
      for (int b = 0; b < NB0; ++b) {
        int c, j;
        bf_coords<L, C, R0, NT>(b, t, cf, c, j);
        const uint32_t o0 = (cbase + (uint32_t)c) * icol + (uint32_t)j * iel;
#pragma unroll
        for (int r = 0; r < R0; ++r) {
          float2 x = ldg2_nt(in, o0 + (uint32_t)(r * (L / R0)) * iel);
          if (a.in_conj) x.y = -x.y;
          dst[b * R0 + r] = x;
        }
      }
    };
    load(col0, v);
    if constexpr (two_tiles_ok(L, C, MODE)) {
      // tiles blk, blk + gridDim.x, ... (< ntiles, at most a.tiles of
      // them): the next tile's loads are in flight while the current one is
      // transformed; the barrier keeps the current tile's last LDS reads
      // ahead of the next one's first writes
      unsigned cur = blk;
#pragma unroll 1
      for (int i = 1; i < a.tiles; ++i) {
        const unsigned nxt = blk + (unsigned)i * gridDim.x;
        if (nxt >= a.ntiles) break;
        float2 v2[P];
        load(nxt * C, v2);
        // the thread index as a fresh value per tile: the tile's index
        // arithmetic is recomputed instead of held across the loop
        int tt = t;
        asm volatile("" : "+v"(tt));
        run_stage<0, L, C, MODE>(v, tw, lds, a, out, cur * C, tt);
        __syncthreads();
#pragma unroll
        for (int p = 0; p < P; ++p) v[p] = v2[p];
        cur = nxt;
      }
      run_stage<0, L, C, MODE>(v, tw, lds, a, out, cur * C, t);
      return;
    }
  }
  run_stage<0, L, C, MODE>(v, tw, lds, a, out, col0, t);
}

// --------------------------------------------------------- R2C untangle
// In place on the M+1 bins of X: thread k in [0, M/2] reads Z[k], Z[M-k] and
// writes X[k], X[M-k] with X[k] = (Zk + Zm*)/2 - i W_n^k (Zk - Zm*)/2.

__global__ void __launch_bounds__(256) k_untangle(float2* __restrict__ X, uint64_t M, int64_t tr_stride,
                                                  RootTable rt) {
  float2* x = X + (int64_t)blockIdx.y * tr_stride;
  const uint64_t half = M / 2;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k <= half; k += stride) {
    const uint64_t km = (k == 0) ? 0 : M - k;
    const float2 zk = x[k];
    const float2 zm = x[km];
    const float2 wk = root(rt, k);
    const float2 xk = untangle_one(zk, zm, wk);
    if (k == 0) {
      // X[M] from Z[0] with W_n^M = -1
      const float2 xm = untangle_one(zk, zk, make_float2(-1.f, 0.f));
      x[0] = xk;
      x[M] = xm;
    } else if (k == half && (M & 1) == 0) {
      x[k] = xk;
    } else {
      // W_n^(M-k) = -conj(W_n^k)
      const float2 wm = make_float2(-wk.x, wk.y);
      const float2 xm = untangle_one(zm, zk, wm);
      x[k] = xk;
      x[km] = xm;
    }
  }
}

// C2R pre-pass, out of place: W[k] = conj( (Xk + Xm*) + i conj(W_n^k) (Xk - Xm*) )
// for k in [0, M), with Xm = X[M-k].  The conjugate feeds the forward passes.
__device__ inline float2 tangle_one(float2 xk, float2 xm, float2 w) {
  const float2 s = make_float2(xk.x + xm.x, xk.y - xm.y);
  const float2 d = make_float2(xk.x - xm.x, xk.y + xm.y);
  const float2 cd = cmul(make_float2(w.x, -w.y), d);
  const float2 z = make_float2(s.x - cd.y, s.y + cd.x);
  return make_float2(z.x, -z.y);
}

__global__ void __launch_bounds__(256) k_tangle(const float2* __restrict__ X, float2* __restrict__ W, uint64_t M,
                                                int64_t in_tr, int64_t out_tr, RootTable rt) {
  const float2* x = X + (int64_t)blockIdx.y * in_tr;
  float2* w = W + (int64_t)blockIdx.y * out_tr;
  const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
  for (uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; k < M; k += stride)
    w[k] = tangle_one(x[k], x[M - k], root(rt, k));
}

// ------------------------------------------------- kernel registry
struct KernelInfo {
  int L, C, mode;
  bool rs;
  void (*fn)(const PassArgs);
  int threads;
  int lds;
  bool two_tiles;  // may run several column tiles per workgroup (PassArgs::tiles)
};
// Column tiles per workgroup of the passes that may take several (the
// 4,096-point row pass at 2^23, 64 trials: 1 tile 0.97 ms, 2 tiles 0.85,
// 4 tiles 0.85, 8 tiles 0.82 per launch).
constexpr int kTilesPerGroup = 8;
// Series size from which the resampling first pass runs tile-major
// (same-box end-to-end rates, trial-major against tile-major: 2^23 27,280
// against 25,960 trials/s, 2^24 12,390 against 11,590; with the default
// long plans 2^25 5,427 against 5,323 and 2^26 2,568 against 2,738).
constexpr uint64_t kTileMajorSeriesBytes = 256ull << 20;

template <int L, int C, int MODE, bool RS>
constexpr KernelInfo kinfo1() {
  return KernelInfo{L, C, MODE, RS, &k_fft_pass<L, C, MODE, RS>, C * L / points_per_thread(L),
                    pass_lds_bytes(L, C, MODE), !RS && two_tiles_ok(L, C, MODE)};
}
// single-pass kernels (C = 1): contiguous, plain and resampling
#define PSX_K1(L) kinfo1<L, 1, PM_CONTIG, false>(), kinfo1<L, 1, PM_CONTIG, true>()
// contiguous rows in and out with column tiles (transposed last passes)
#define PSX_KR(L, C) kinfo1<L, C, PM_CONTIG, false>()
// column-tile kernels: strided in/out with twiddle (plain and resampling)
// and rows-to-columns
#define PSX_KC(L, C) \
  kinfo1<L, C, PM_COLS_TW, false>(), kinfo1<L, C, PM_COLS_TW, true>(), kinfo1<L, C, PM_ROWS_TO_COLS, false>()

const KernelInfo kKernels[] = {
    PSX_K1(8), PSX_K1(16), PSX_K1(32), PSX_K1(64), PSX_K1(128), PSX_K1(256), PSX_K1(512), PSX_K1(1024),
    PSX_K1(2048), PSX_K1(4096), PSX_K1(6), PSX_K1(12), PSX_K1(24), PSX_K1(48), PSX_K1(96), PSX_K1(192),
    PSX_K1(384), PSX_K1(768), PSX_K1(1536), PSX_K1(3072),
    PSX_KC(32, 16), PSX_KC(64, 16), PSX_KC(128, 16), PSX_KC(256, 16), PSX_KC(512, 16), PSX_KC(1024, 16),
    PSX_KC(64, 128), PSX_KC(128, 128), PSX_KC(64, 64), PSX_KC(128, 64), PSX_KC(64, 32), PSX_KC(128, 32),
    PSX_KC(64, 8), PSX_KC(128, 8), PSX_KC(256, 8), PSX_KC(512, 8), PSX_KC(1024, 8), PSX_KC(2048, 8),
    PSX_KC(256, 4), PSX_KC(512, 4), PSX_KC(1024, 4), PSX_KC(2048, 4), PSX_KC(4096, 4),
    PSX_KC(1024, 2), PSX_KC(2048, 2), PSX_KC(4096, 2),
    PSX_KC(48, 16), PSX_KC(96, 16), PSX_KC(192, 16), PSX_KC(384, 16), PSX_KC(768, 16),
    PSX_KC(96, 8), PSX_KC(192, 8), PSX_KC(384, 8), PSX_KC(768, 8), PSX_KC(1536, 8),
    PSX_KC(384, 4), PSX_KC(768, 4), PSX_KC(1536, 4), PSX_KC(3072, 4),
    PSX_KR(128, 16), PSX_KR(256, 16), PSX_KR(512, 16),
    PSX_KR(512, 8), PSX_KR(1024, 8), PSX_KR(2048, 8), PSX_KR(768, 8), PSX_KR(1536, 8),
    PSX_KR(1024, 4), PSX_KR(2048, 4), PSX_KR(4096, 4), PSX_KR(1536, 4), PSX_KR(3072, 4),
    PSX_KR(2048, 2), PSX_KR(4096, 2),
};

const KernelInfo* find_kernel(int L, int C, int mode = -1, bool rs = false) {
  for (const KernelInfo& k : kKernels)
    if (k.L == L && k.C == C && (mode < 0 || (k.mode == mode && k.rs == rs))) return &k;
  return nullptr;
}

bool g_lds_configured = false;

int configure_lds() {
  if (g_lds_configured) return PSX_OK;
  for (const KernelInfo& k : kKernels) {
    if (k.lds > 65536) {
      hipError_t e = hipFuncSetAttribute((const void*)k.fn, hipFuncAttributeMaxDynamicSharedMemorySize, k.lds);
      if (e != hipSuccess)
        return psx::fail(PSX_ERR_HIP, "hipFuncSetAttribute(L=%d,C=%d,lds=%d): %s", k.L, k.C, k.lds,
                         hipGetErrorString(e));
    }
  }
  g_lds_configured = true;
  return PSX_OK;
}

// ------------------------------------------------------- host tables
void root_table_host(uint64_t m, uint32_t shift, std::vector<float2>& lo, std::vector<float2>& hi) {
  // W_m^x = W_m^(hi part) * W_m^(lo part), x = (h << shift) + l
  const uint64_t T = 1ull << shift;
  lo.resize(T);
  hi.resize((m + T - 1) / T + 1);
  const long double tp = -2.0L * 3.14159265358979323846264338327950288L / (long double)m;
  for (uint64_t l = 0; l < T; ++l) {
    long double a = tp * (long double)l;
    lo[l] = make_float2((float)cosl(a), (float)sinl(a));
  }
  for (uint64_t h = 0; h < hi.size(); ++h) {
    long double a = tp * (long double)(h * T);
    hi[h] = make_float2((float)cosl(a), (float)sinl(a));
  }
}

void twl_host(int L, std::vector<float2>& t) {
  t.resize(L);
  const long double tp = -2.0L * 3.14159265358979323846264338327950288L / (long double)L;
  for (int e = 0; e < L; ++e) t[e] = make_float2((float)cosl(tp * e), (float)sinl(tp * e));
}

}  // namespace

// =================================================== plan
namespace psx {
namespace fft {

bool valid_complex_len(uint64_t m) {
  if (m < 3) return false;
  uint64_t p = m % 3 == 0 ? m / 3 : m;
  return (p & (p - 1)) == 0;
}

static bool len_supported(uint64_t L, bool multi) {
  for (const KernelInfo& k : kKernels)
    if ((uint64_t)k.L == L && (multi ? k.C > 1 : k.C == 1)) return true;
  return false;
}

// Default decomposition: one pass when a single-workgroup kernel exists,
// else the two-pass split with the longest contiguous (last) pass, whose
// rows are read and written whole, and the shortest strided pass (measured
// fastest on MI355X: 1024 x 4096 beats 2048 x 2048 for 2^22 points) while
// that strided pass is at most 2,048 points, else three passes as below.
bool default_decomp(uint64_t M, Decomp& d) {
  d = Decomp();
  if (!valid_complex_len(M)) return false;
  if (len_supported(M, false)) {
    d.npass = 1;
    d.L[0] = (int)M;
    d.C[0] = 1;
    return true;
  }
  uint64_t bM1 = 0, bM2 = 0;
  for (uint64_t M1 = 32; M1 <= 4096 && M1 < M; M1 = (M1 % 3 == 0) ? (M1 / 3) * 4 : M1 + M1 / 2) {
    if (M % M1) continue;
    const uint64_t M2 = M / M1;
    if (!len_supported(M1, true) || !len_supported(M2, true)) continue;
    if (M1 > bM1) {
      bM1 = M1;
      bM2 = M2;
    }
  }
  auto two = [&]() {
    d.npass = 2;
    d.L[0] = (int)bM2;
    d.L[1] = (int)bM1;
    return pick_tiles(d);
  };
  // two passes while the strided first pass is at most 2,048 points (8+
  // column tiles, 64-byte runs); a 4,096-point strided pass runs 4-column
  // tiles, 32-byte runs, and three passes with a long contiguous last pass
  // are faster (2^25 samples: 4,096 x 4,096 3,000 trials/s, 64 x 64 x 4,096
  // 5,090, same box)
  if (bM1 && bM2 <= 2048) return two();
  // three passes: the longest contiguous last pass (row-layout kernels up to
  // 4,096 points) and two short strided passes of at most 512 points, as
  // even as possible with the first (resampling) pass the longer one
  // (2^26 samples: 128 x 64 x 4,096 2,517 trials/s, 512 x 128 x 512 2,343)
  for (uint64_t M1 : {4096ull, 3072ull, 2048ull, 1536ull, 1024ull}) {
    if (M % M1 || !len_supported(M1, true)) continue;
    const uint64_t rest = M / M1;
    uint64_t bq = 0, bm = 0;
    for (uint64_t M2 = 32; M2 <= 512; M2 = (M2 % 3 == 0) ? (M2 / 3) * 4 : M2 + M2 / 2) {
      if (rest % M2 || !len_supported(M2, true)) continue;
      const uint64_t Q = rest / M2;
      if (Q < M2 || Q > 512 || !len_supported(Q, true)) continue;
      if (!bq || Q < bq) {
        bq = Q;
        bm = M2;
      }
    }
    if (bq) {
      d.npass = 3;
      d.L[0] = (int)bq;
      d.L[1] = (int)bm;
      d.L[2] = (int)M1;
      return pick_tiles(d);
    }
  }
  if (bM1) return two();
  // otherwise the three passes whose longest is shortest
  uint64_t best = 0, bQ = 0;
  for (uint64_t M1 = 32; M1 <= 4096; M1 = (M1 % 3 == 0) ? (M1 / 3) * 4 : M1 + M1 / 2) {
    if (M % M1 || !len_supported(M1, true)) continue;
    for (uint64_t M2 = 32; M2 <= 4096; M2 = (M2 % 3 == 0) ? (M2 / 3) * 4 : M2 + M2 / 2) {
      if ((M / M1) % M2 || !len_supported(M2, true)) continue;
      const uint64_t Q = M / M1 / M2;
      if (!len_supported(Q, true)) continue;
      uint64_t score = M1;
      if (M2 > score) score = M2;
      if (Q > score) score = Q;
      if (!best || score < best || (score == best && M1 > bM1)) {
        best = score;
        bM1 = M1;
        bM2 = M2;
        bQ = Q;
      }
    }
  }
  if (!best) return false;
  d.npass = 3;
  d.L[0] = (int)bQ;
  d.L[1] = (int)bM2;
  d.L[2] = (int)bM1;
  return pick_tiles(d);
}

// Widest column tile whose LDS fits 80 KiB (two workgroups per CU) for the
// strided passes; the last pass, which reads and writes whole rows, may use
// a CU's whole LDS.  The short strided passes of three-pass plans take the
// widest tiles that exist for them (64 and 128 points: 128 columns, 1 KB
// runs; same-box end to end at 2^26 with 16 / 32 / 64 / 128 columns 2,569 /
// 2,606 / 2,695 / 2,715 trials/s, at 2^25 5,101 / 5,208 / 5,250 / 5,318,
// same candidates).
bool pick_tiles(Decomp& d) {
  for (int p = 0; p < d.npass; ++p) {
    if (d.C[p] > 0 && find_kernel(d.L[p], d.C[p])) continue;
    const int pref3[] = {128, 64, 32, 16, 8, 4, 2}, pref2[] = {16, 8, 4, 2};
    const bool last = (p == d.npass - 1 && d.npass > 1);
    const int* pref = (d.npass == 3 && !last) ? pref3 : pref2;
    const int npref = (d.npass == 3 && !last) ? 7 : 4;
    const uint32_t cap = last ? 160u * 1024u : 80u * 1024u;
    const int mode = last ? PM_ROWS_TO_COLS : PM_COLS_TW;
    d.C[p] = 0;
    for (int i = 0; i < npref; ++i) {
      const KernelInfo* k = find_kernel(d.L[p], pref[i], mode, false);
      if (k && (uint32_t)k->lds <= cap) {
        d.C[p] = pref[i];
        break;
      }
    }
    if (!d.C[p])
      for (int i = 0; i < npref; ++i)
        if (find_kernel(d.L[p], pref[i])) {
          d.C[p] = pref[i];
          break;
        }
    if (!d.C[p]) return false;
  }
  return true;
}

bool decomp_supported(const Decomp& d) {
  for (int p = 0; p < d.npass; ++p)
    if (!find_kernel(d.L[p], d.C[p])) return false;
  return true;
}

}  // namespace fft
}  // namespace psx

struct PassLaunch {
  const KernelInfo* k;
  const KernelInfo* krs;  // resampling variant of a first pass (batched search)
  PassArgs a;           // in/out pointers filled at exec time from src/dst
  int src, dst;         // 0 = user input, 1 = work, 2 = user output
  unsigned gx, gy, gz;
};

struct psx_fft_plan {
  uint64_t n = 0, M = 0, nb = 0;
  int kind = 0;
  uint32_t batch = 1;
  psx::fft::Decomp d;
  std::vector<PassLaunch> passes;
  // Multi-pass R2C plans whose last pass has a contiguous-row kernel also
  // carry a version of that pass writing contiguous rows: Y[r*L + e] =
  // Z[r + R*e] (L the last pass's length, R = M/L), the transposed layout
  // the batched spectrum pass reads.
  PassLaunch rows_last{};
  bool has_rows_last = false;
  float2* tables = nullptr;  // all device tables in one allocation
  float2* work = nullptr;    // M complex per transform (when needed)
  RootTable rt{};            // W_n roots for (un)tangle
};

namespace {

int build_passes(psx_fft_plan* p, bool device_tables) {
  using psx::fft::Decomp;
  const Decomp& d = p->d;
  const uint64_t M = p->M;
  // --- device tables: per-pass W_L, four-step roots, W_n roots
  std::vector<float2> host;
  std::vector<size_t> twl_off(d.npass), tw4lo_off(d.npass), tw4hi_off(d.npass), tw4r_off(d.npass);
  std::vector<uint32_t> tw4_shift(d.npass);
  auto append = [&](const std::vector<float2>& v) {
    size_t o = host.size();
    host.insert(host.end(), v.begin(), v.end());
    while (host.size() % 32) host.push_back(make_float2(0.f, 0.f));
    return o;
  };
  std::vector<float2> tmp, lo, hi;
  // four-step twiddle sizes per pass
  uint64_t mtw[3] = {0, 0, 0};
  if (d.npass == 2) mtw[0] = M;
  if (d.npass == 3) {
    mtw[0] = M;
    mtw[1] = (uint64_t)d.L[1] * (uint64_t)d.L[2];
  }
  for (int q = 0; q < d.npass; ++q) {
    twl_host(d.L[q], tmp);
    twl_off[q] = append(tmp);
    if (mtw[q]) {
      uint32_t sh = 0;
      while ((1ull << (sh + 1)) * (1ull << (sh + 1)) <= mtw[q] && (mtw[q] % (1ull << (sh + 1))) == 0) ++sh;
      tw4_shift[q] = sh;
      root_table_host(mtw[q], sh, lo, hi);
      tw4lo_off[q] = append(lo);
      tw4hi_off[q] = append(hi);
      // W_T^k, T = Mtw * R / L with R the pass's last radix (the output
      // twiddle's column factor)
      const uint64_t T = mtw[q] * (uint64_t)psx::fft::radix_at(d.L[q], psx::fft::num_stages(d.L[q]) - 1) /
                         (uint64_t)d.L[q];
      tmp.resize(T);
      const long double tpr = -2.0L * 3.14159265358979323846264338327950288L / (long double)T;
      for (uint64_t k = 0; k < T; ++k) tmp[k] = make_float2((float)cosl(tpr * k), (float)sinl(tpr * k));
      tw4r_off[q] = append(tmp);
    }
  }
  uint32_t rsh = 0;
  const uint64_t n = p->n;
  while ((1ull << (rsh + 1)) * (1ull << (rsh + 1)) <= n && (n % (1ull << (rsh + 1))) == 0) ++rsh;
  root_table_host(n, rsh, lo, hi);
  const size_t rlo = append(lo), rhi = append(hi);
  if (device_tables) {
    PSX_CHECK_HIP(hipMalloc(&p->tables, host.size() * sizeof(float2)));
    PSX_CHECK_HIP(hipMemcpy(p->tables, host.data(), host.size() * sizeof(float2), hipMemcpyHostToDevice));
  }
  p->rt = RootTable{p->tables + rlo, p->tables + rhi, rsh, (uint32_t)((1ull << rsh) - 1)};

  // --- passes
  const bool c2r = p->kind == PSX_FFT_C2R;
  const int64_t trM = (int64_t)M;                // complex stride of work / real-as-complex buffers
  const int64_t trNB = (int64_t)(M + 1);         // complex stride of spectra
  const int64_t in_tr = c2r ? trM : trM;         // passes read z (R2C: input as complex) or work
  auto make = [&](int q, int src, int dst, int mode) {
    PassLaunch L{};
    L.k = find_kernel(d.L[q], d.C[q], mode, false);
    L.krs = find_kernel(d.L[q], d.C[q], mode, true);
    L.src = src;
    L.dst = dst;
    PassArgs& a = L.a;
    a.twL = p->tables + twl_off[q];
    if (mtw[q]) {
      a.tw4 = 1;
      a.tw4_lo = p->tables + tw4lo_off[q];
      a.tw4_hi = p->tables + tw4hi_off[q];
      a.tw4_r = p->tables + tw4r_off[q];
      a.tw4_shift = tw4_shift[q];
      a.tw4_mask = (uint32_t)((1ull << tw4_shift[q]) - 1);
    }
    a.in_tr = (src == 2) ? trNB : in_tr;
    a.out_tr = (dst == 2) ? (c2r ? trM : trNB) : trM;
    return L;
  };
  if (d.npass == 1) {
    PassLaunch L = make(0, c2r ? 1 : 0, 2, PM_CONTIG);
    L.a.in_col = 0; L.a.in_elem = 1; L.a.in_sub = 0;
    L.a.out_col = 0; L.a.out_elem = 1; L.a.out_sub = 0;
    L.gx = 1; L.gy = 1; L.gz = p->batch;
    p->passes.push_back(L);
  } else if (d.npass == 2) {
    const int64_t M2 = d.L[0], M1 = d.L[1];
    PassLaunch A = make(0, c2r ? 1 : 0, 1, PM_COLS_TW);
    A.a.in_col = 1; A.a.in_elem = M1; A.a.in_sub = 0;
    A.a.out_col = 1; A.a.out_elem = M1; A.a.out_sub = 0;
    A.a.in_cf = 1; A.a.out_cf = 1;
    A.gx = (unsigned)(M1 / d.C[0]); A.gy = 1; A.gz = p->batch;
    PassLaunch B = make(1, 1, 2, PM_ROWS_TO_COLS);
    B.a.in_col = M1; B.a.in_elem = 1; B.a.in_sub = 0;
    B.a.out_col = 1; B.a.out_elem = M2; B.a.out_sub = 0;
    B.a.in_cf = 0; B.a.out_cf = 1;
    B.gx = (unsigned)(M2 / d.C[1]); B.gy = 1; B.gz = p->batch;
    p->passes.push_back(A);
    p->passes.push_back(B);
    if (!c2r) {
      PassLaunch Bt = make(1, 1, 2, PM_CONTIG);
      Bt.a.in_col = M1; Bt.a.in_elem = 1; Bt.a.in_sub = 0;
      Bt.a.out_col = M1; Bt.a.out_elem = 1; Bt.a.out_sub = 0;
      Bt.a.in_cf = 0; Bt.a.out_cf = 0;
      Bt.gx = B.gx; Bt.gy = 1; Bt.gz = p->batch;
      if (Bt.k) {
        p->rows_last = Bt;
        p->has_rows_last = true;
      }
    }
  } else {
    const int64_t Q = d.L[0], M2 = d.L[1], M1 = d.L[2], P = M1 * M2;
    PassLaunch A = make(0, c2r ? 1 : 0, 1, PM_COLS_TW);
    A.a.in_col = 1; A.a.in_elem = P; A.a.in_sub = 0;
    A.a.out_col = 1; A.a.out_elem = P; A.a.out_sub = 0;
    A.a.in_cf = 1; A.a.out_cf = 1;
    A.gx = (unsigned)(P / d.C[0]); A.gy = 1; A.gz = p->batch;
    PassLaunch B = make(1, 1, 1, PM_COLS_TW);
    B.a.in_col = 1; B.a.in_elem = M1; B.a.in_sub = P;
    B.a.out_col = 1; B.a.out_elem = M1; B.a.out_sub = P;
    B.a.in_cf = 1; B.a.out_cf = 1;
    B.gx = (unsigned)(M1 / d.C[1]); B.gy = (unsigned)Q; B.gz = p->batch;
    PassLaunch Cc = make(2, 1, 2, PM_ROWS_TO_COLS);
    Cc.a.in_col = P; Cc.a.in_elem = 1; Cc.a.in_sub = M1;
    Cc.a.out_col = 1; Cc.a.out_elem = Q * M2; Cc.a.out_sub = Q;
    Cc.a.in_cf = 0; Cc.a.out_cf = 1;
    Cc.gx = (unsigned)(Q / d.C[2]); Cc.gy = (unsigned)M2; Cc.gz = p->batch;
    p->passes.push_back(A);
    p->passes.push_back(B);
    p->passes.push_back(Cc);
    if (!c2r) {
      // the same last pass writing contiguous rows: transform (c, y) to row
      // c + Q*y of M1 elements, i.e. row r holds Z[r + (Q*M2)*e]
      PassLaunch Ct = make(2, 1, 2, PM_CONTIG);
      Ct.a.in_col = P; Ct.a.in_elem = 1; Ct.a.in_sub = M1;
      Ct.a.out_col = M1; Ct.a.out_elem = 1; Ct.a.out_sub = Q * M1;
      Ct.a.in_cf = 0; Ct.a.out_cf = 0;
      Ct.gx = Cc.gx; Ct.gy = Cc.gy; Ct.gz = p->batch;
      if (Ct.k) {
        p->rows_last = Ct;
        p->has_rows_last = true;
      }
    }
  }
  if (c2r) p->passes.back().a.out_conj = 1;
  for (const PassLaunch& L : p->passes) {
    if (!L.k) return psx::fail(PSX_ERR_UNSUPPORTED, "no FFT kernel for this decomposition");
    if (L.gx < 1 || L.gy > 65535 || L.gz > 65535) return psx::fail(PSX_ERR_UNSUPPORTED, "FFT grid too large");
  }
  return PSX_OK;
}

}  // namespace

namespace psx {
namespace fft {

int plan_create_with(uint64_t n, int kind, uint32_t batch, const Decomp* forced, psx_fft_plan** out) {
  PSX_REQUIRE(out, "psx_fft_plan_create: null plan");
  *out = nullptr;
  PSX_REQUIRE(kind == PSX_FFT_R2C || kind == PSX_FFT_C2R, "psx_fft_plan_create: bad kind %d", kind);
  PSX_REQUIRE(n >= 16 && n <= (1ull << 28) && (n % 2) == 0 && valid_complex_len(n / 2),
              "psx_fft_plan_create: unsupported length %llu (need 2^k or 3*2^k, 16 <= n <= 2^28)",
              (unsigned long long)n);
  PSX_REQUIRE(batch >= 1 && batch <= 65535, "psx_fft_plan_create: batch must be 1..65535");
  int rc = configure_lds();
  if (rc) return rc;
  psx_fft_plan* p = new psx_fft_plan;
  p->n = n;
  p->M = n / 2;
  p->nb = n / 2 + 1;
  p->kind = kind;
  p->batch = batch;
  bool okd = forced ? (p->d = *forced, pick_tiles(p->d)) : default_decomp(p->M, p->d);
  if (!okd || !decomp_supported(p->d)) {
    delete p;
    return psx::fail(PSX_ERR_UNSUPPORTED, "psx_fft_plan_create: no decomposition for n=%llu",
                     (unsigned long long)n);
  }
  uint64_t prod = 1;
  for (int q = 0; q < p->d.npass; ++q) prod *= (uint64_t)p->d.L[q];
  if (prod != p->M) {
    delete p;
    return psx::fail(PSX_ERR_ARG, "psx_fft_plan_create: decomposition does not multiply to M");
  }
  rc = build_passes(p, true);
  if (!rc) {
    const bool need_work = (kind == PSX_FFT_C2R) || p->d.npass > 1;
    if (need_work) {
      hipError_t e = hipMalloc(&p->work, sizeof(float2) * p->M * batch);
      if (e != hipSuccess) rc = psx::fail(PSX_ERR_ALLOC, "FFT work buffer: %s", hipGetErrorString(e));
    }
  }
  if (rc) {
    psx_fft_plan_destroy(p);
    return rc;
  }
  *out = p;
  return psx::ok();
}

int check_pass_bounds(const PassLaunch& L, const PassArgs& a, uint32_t gz, uint64_t cap_in, uint64_t cap_out,
                      uint64_t m_lim);

int plan_exec(psx_fft_plan* p, const void* in, void* out, hipStream_t hs) {
  const unsigned tb = 256;
  if (p->kind == PSX_FFT_C2R) {
    const unsigned g = psx::grid_for(p->M, tb);
    k_tangle<<<dim3(g, p->batch), tb, 0, hs>>>((const float2*)in, p->work, p->M, (int64_t)p->nb, (int64_t)p->M,
                                               p->rt);
    PSX_CHECK_LAUNCH("k_tangle");
  }
  for (PassLaunch& L : p->passes) {
    PassArgs a = L.a;
    a.in = L.src == 0 ? (const float2*)in : (L.src == 1 ? p->work : (const float2*)out);
    a.out = L.dst == 2 ? (float2*)out : p->work;
    // capacities in complex elements: user input (R2C) n/2 per transform,
    // work n/2, user output n/2+1 (R2C) or n/2 (C2R) per transform
    const uint64_t cap_in = p->M * p->batch;
    const uint64_t cap_out = L.dst == 2 ? (p->kind == PSX_FFT_R2C ? p->nb : p->M) * p->batch : p->M * p->batch;
    int rc = check_pass_bounds(L, a, L.gz, cap_in, cap_out, 0);
    if (rc) return rc;
    hipLaunchKernelGGL(L.k->fn, dim3(L.gx, L.gy, L.gz), dim3(L.k->threads), L.k->lds, hs, a);
    PSX_CHECK_LAUNCH("k_fft_pass");
  }
  if (p->kind == PSX_FFT_R2C) {
    const unsigned g = psx::grid_for(p->M / 2 + 1, tb);
    k_untangle<<<dim3(g, p->batch), tb, 0, hs>>>((float2*)out, p->M, (int64_t)p->nb, p->rt);
    PSX_CHECK_LAUNCH("k_untangle");
  }
  return PSX_OK;
}

// Largest element index a pass launch reads and writes (all strides are
// non-negative), checked against the capacities of its buffers before the
// launch.  cap_in == 0 skips the input check; m_lim > 0 checks the complex
// index of a resampling first pass (m < m_lim keeps 2m+1 < n).
int check_pass_bounds(const PassLaunch& L, const PassArgs& a, uint32_t gz, uint64_t cap_in, uint64_t cap_out,
                      uint64_t m_lim) {
  const uint64_t C = (uint64_t)L.k->C, Ln = (uint64_t)L.k->L;
  const uint64_t cols = (uint64_t)L.gx * C - 1, gy = (uint64_t)L.gy - 1, gzz = (uint64_t)gz - 1;
  if (a.in_col < 0 || a.in_elem < 0 || a.in_sub < 0 || a.in_tr < 0 || a.out_col < 0 || a.out_elem < 0 ||
      a.out_sub < 0 || a.out_tr < 0)
    return psx::fail(PSX_ERR_ARG, "FFT pass: negative stride");
  const uint64_t in_max = gzz * a.in_tr + gy * a.in_sub + cols * a.in_col + (Ln - 1) * a.in_elem;
  const uint64_t out_max = gzz * a.out_tr + gy * a.out_sub + cols * a.out_col + (Ln - 1) * a.out_elem;
  const uint64_t m_max = gy * a.in_sub + cols * a.in_col + (Ln - 1) * a.in_elem;
  if (cap_in && in_max >= cap_in)
    return psx::fail(PSX_ERR_ARG, "FFT pass (L=%d,C=%d) would read element %llu of %llu", L.k->L, L.k->C,
                     (unsigned long long)in_max, (unsigned long long)cap_in);
  if (out_max >= cap_out)
    return psx::fail(PSX_ERR_ARG, "FFT pass (L=%d,C=%d) would write element %llu of %llu", L.k->L, L.k->C,
                     (unsigned long long)out_max, (unsigned long long)cap_out);
  if (m_lim && m_max >= m_lim)
    return psx::fail(PSX_ERR_ARG, "FFT pass (L=%d,C=%d) would resample beyond the series", L.k->L, L.k->C);
  return PSX_OK;
}

int exec_fused_pass(psx_fft_plan* p, int q, const TrialDesc* d_descs, uint32_t ntrials, float2* zout, hipStream_t hs,
                    bool rows_out) {
  if (p->kind != PSX_FFT_R2C || ntrials < 1 || ntrials > p->batch || q < 0 || q >= (int)p->passes.size())
    return psx::fail(PSX_ERR_ARG, "exec_fused_pass: bad plan, pass or trial count");
  const bool last = q == (int)p->passes.size() - 1;
  if (rows_out && last && !p->has_rows_last)
    return psx::fail(PSX_ERR_UNSUPPORTED, "exec_fused_pass: plan has no row-layout last pass");
  PassLaunch& L = (rows_out && last) ? p->rows_last : p->passes[q];
  PassArgs a = L.a;
  a.in = L.src == 1 ? p->work : nullptr;
  a.out = L.dst == 2 ? zout : p->work;
  // zout holds the half-length spectra Z of the trials back to back (stride
  // M), not the M+1-bin R2C output layout of the plain plan.
  if (L.dst == 2) a.out_tr = (int64_t)p->M;
  if (q == 0) {
    a.resample = 1;
    a.rs_n = p->n;
    a.descs = d_descs;
    a.rs_tile_major = p->n * sizeof(float) >= kTileMajorSeriesBytes ? 1 : 0;
  }
  const uint64_t cap_in = p->M * p->batch, cap_out = L.dst == 2 ? p->M * p->batch : p->M * p->batch;
  int rc = check_pass_bounds(L, a, ntrials, q == 0 ? 0 : cap_in, cap_out, q == 0 ? p->M : 0);
  if (rc) return rc;
  const KernelInfo* k = q == 0 ? L.krs : L.k;
  if (!k) return psx::fail(PSX_ERR_UNSUPPORTED, "exec_fused_pass: no resampling kernel for this pass");
  // the resampling kernel orders its blocks tile-major over x and z only
  if (q == 0 && L.gy != 1) return psx::fail(PSX_ERR_UNSUPPORTED, "exec_fused_pass: first pass with sub-transforms");
  unsigned gx = L.gx;
  if (k->two_tiles && gx > 1) {
    a.tiles = kTilesPerGroup;
    a.ntiles = gx;
    gx = (gx + kTilesPerGroup - 1) / kTilesPerGroup;
  }
  hipLaunchKernelGGL(k->fn, dim3(gx, L.gy, ntrials), dim3(k->threads), k->lds, hs, a);
  PSX_CHECK_LAUNCH("k_fft_pass(fused)");
  return PSX_OK;
}

int exec_fused_r2c(psx_fft_plan* p, const TrialDesc* d_descs, uint32_t ntrials, float2* zout, hipStream_t hs,
                   bool rows_out) {
  for (int q = 0; q < (int)p->passes.size(); ++q) {
    int rc = exec_fused_pass(p, q, d_descs, ntrials, zout, hs, rows_out);
    if (rc) return rc;
  }
  return PSX_OK;
}

bool rows_layout(const psx_fft_plan* p, uint32_t* rows, uint32_t* row_len) {
  if (!p->has_rows_last) return false;
  const int last = p->d.npass - 1;
  uint32_t r = 1;
  for (int q = 0; q < last; ++q) r *= (uint32_t)p->d.L[q];
  *rows = r;
  *row_len = (uint32_t)p->d.L[last];
  return true;
}

int num_passes(const psx_fft_plan* p) { return (int)p->passes.size(); }
bool pass_in_place(const psx_fft_plan* p, int q) {
  return q >= 0 && q < (int)p->passes.size() && p->passes[q].src == 1 && p->passes[q].dst == 1;
}

RootTable root_table(const psx_fft_plan* p) { return p->rt; }
uint64_t complex_len(const psx_fft_plan* p) { return p->M; }
const Decomp& plan_decomp(const psx_fft_plan* p) { return p->d; }

}  // namespace fft
}  // namespace psx

extern "C" int psx_fft_plan_create(uint64_t n, int kind, uint32_t batch, psx_fft_plan** plan) {
  return psx::fft::plan_create_with(n, kind, batch, nullptr, plan);
}

extern "C" int psx_fft_plan_destroy(psx_fft_plan* plan) {
  if (!plan) return psx::ok();
  if (plan->tables) (void)hipFree(plan->tables);
  if (plan->work) (void)hipFree(plan->work);
  delete plan;
  return psx::ok();
}

extern "C" int psx_fft_exec(psx_fft_plan* plan, const void* in, void* out, psx_stream s) {
  PSX_REQUIRE(plan && in && out, "psx_fft_exec: null argument");
  PSX_REQUIRE(in != out, "psx_fft_exec: in-place transforms are not supported");
  PSX_REQUIRE(((uintptr_t)in & 7) == 0 && ((uintptr_t)out & 7) == 0, "psx_fft_exec: buffers must be 8-byte aligned");
  int rc = psx::fft::plan_exec(plan, in, out, as_stream(s));
  if (rc) return rc;
  return psx::ok();
}

extern "C" int psx_fft_plan_describe(const psx_fft_plan* p, char* buf, size_t cap) {
  PSX_REQUIRE(p && buf && cap, "psx_fft_plan_describe: bad arguments");
  std::string s = "n=" + std::to_string(p->n) + (p->kind == PSX_FFT_R2C ? " R2C" : " C2R") +
                  " batch=" + std::to_string(p->batch) + " passes=";
  for (int q = 0; q < p->d.npass; ++q)
    s += (q ? "x" : "") + std::to_string(p->d.L[q]) + "/C" + std::to_string(p->d.C[q]);
  s += " launches=" + std::to_string(p->passes.size() + 1);
  strncpy(buf, s.c_str(), cap - 1);
  buf[cap - 1] = 0;
  return psx::ok();
}

// Host-only geometry check of a plan: builds the pass launches without device
// tables and validates every pass's largest read and write index against its
// buffers, for the plain transform and for the fused batched use.
extern "C" int psx_fft_plan_create_decomp(uint64_t n, int kind, uint32_t batch, int npass, const int* lens,
                                          const int* cols, psx_fft_plan** plan) {
  PSX_REQUIRE(npass >= 1 && npass <= 3 && lens && cols, "psx_fft_plan_create_decomp: bad decomposition");
  psx::fft::Decomp d;
  d.npass = npass;
  for (int q = 0; q < npass; ++q) {
    d.L[q] = lens[q];
    d.C[q] = cols[q];
  }
  return psx::fft::plan_create_with(n, kind, batch, &d, plan);
}

extern "C" int psx_fft_layout_check(uint64_t n, int kind, uint32_t batch, int npass, const int* lens, const int* cols) {
  PSX_REQUIRE(n >= 16 && (n % 2) == 0 && psx::fft::valid_complex_len(n / 2), "psx_fft_layout_check: bad length");
  PSX_REQUIRE(kind == PSX_FFT_R2C || kind == PSX_FFT_C2R, "psx_fft_layout_check: bad kind");
  PSX_REQUIRE(batch >= 1 && batch <= 65535, "psx_fft_layout_check: bad batch");
  psx_fft_plan p;
  p.n = n;
  p.M = n / 2;
  p.nb = n / 2 + 1;
  p.kind = kind;
  p.batch = batch;
  bool okd;
  if (npass > 0) {
    PSX_REQUIRE(npass <= 3 && lens && cols, "psx_fft_layout_check: bad decomposition");
    p.d.npass = npass;
    for (int q = 0; q < npass; ++q) {
      p.d.L[q] = lens[q];
      p.d.C[q] = cols[q];
    }
    okd = psx::fft::pick_tiles(p.d);
  } else {
    okd = psx::fft::default_decomp(p.M, p.d);
  }
  if (!okd || !psx::fft::decomp_supported(p.d)) return psx::fail(PSX_ERR_UNSUPPORTED, "no decomposition");
  uint64_t prod = 1;
  for (int q = 0; q < p.d.npass; ++q) prod *= (uint64_t)p.d.L[q];
  PSX_REQUIRE(prod == p.M, "psx_fft_layout_check: decomposition does not multiply to n/2");
  int rc = build_passes(&p, false);
  if (rc) return rc;
  for (size_t q = 0; q < p.passes.size(); ++q) {
    const PassLaunch& L = p.passes[q];
    const uint64_t cap_in = p.M * p.batch;
    const uint64_t cap_out = L.dst == 2 ? (kind == PSX_FFT_R2C ? p.nb : p.M) * p.batch : p.M * p.batch;
    rc = psx::fft::check_pass_bounds(L, L.a, L.gz, cap_in, cap_out, 0);
    if (rc) return rc;
    if (kind == PSX_FFT_R2C) {
      PassArgs a = L.a;
      if (L.dst == 2) a.out_tr = (int64_t)p.M;
      rc = psx::fft::check_pass_bounds(L, a, p.batch, q == 0 ? 0 : cap_in, p.M * p.batch, q == 0 ? p.M : 0);
      if (rc) return rc;
    }
  }
  if (kind == PSX_FFT_R2C && p.has_rows_last) {
    // the fused search's row-layout last pass (half-spectra back to back)
    PassArgs a = p.rows_last.a;
    a.out_tr = (int64_t)p.M;
    const uint64_t cap = p.M * p.batch;
    rc = psx::fft::check_pass_bounds(p.rows_last, a, p.batch, p.passes.size() == 1 ? 0 : cap, cap, 0);
    if (rc) return rc;
  }
  return psx::ok();
}
