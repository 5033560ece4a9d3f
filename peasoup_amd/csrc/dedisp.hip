// dedisp.hip - incoherent dedispersion of SIGPROC filterbank data (gfx950).
//
// The reference dedisperses with libdedisp (dedisp_execute through
// Dedisperser, include/transforms/dedisperser.hpp:105-186), a dependency not
// in its tree.  This file restates libdedisp's direct algorithm with its
// published delay table, DM-list recurrence and float output scaling:
//   delay(dm, c) = rn(dm * table[c]),
//   table[c]     = 4.15e3 / tsamp * (1/(f0 + c df)^2 - 1/f0^2)   (float)
//   out[dm][t]   = sum_c v(t + delay(dm, c), c) / ((2^nbits - 1) * nchans)
// Two kernels:
//   k_unpack_t  packed time-major samples -> one u8 row per channel
//               (a 64-sample x 64-byte tile transposed through LDS);
//   k_dedisp    one block per (DM, 1024-sample tile): the channel loop reads
//               each row at the DM's delay with consecutive lanes on
//               consecutive samples; consecutive blocks are neighbouring DMs
//               of the same tile, so their rows are shared in L2.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "trace.hpp"
#include "psx.h"

namespace {

constexpr unsigned kB = 256;
constexpr unsigned kTT = 64;   // samples per unpack tile
constexpr unsigned kTBy = 64;  // bytes per unpack tile
constexpr unsigned kQ = 4;     // samples per thread in k_dedisp
constexpr uint32_t kMaxChans = 8192;  // the per-DM delays live in 32 KiB of LDS

template <int NBITS>
__global__ void __launch_bounds__(kB) k_unpack_t(const uint8_t* __restrict__ in, uint64_t nsamps, uint32_t nchans,
                                                 uint8_t* __restrict__ out) {
  constexpr unsigned K = 8 / NBITS;  // channels per byte
  constexpr unsigned mask = (1u << NBITS) - 1u;
  __shared__ uint8_t tile[kTT][kTBy + 4];
  const uint64_t rowbytes = (uint64_t)nchans * NBITS / 8;
  const uint64_t t0 = (uint64_t)blockIdx.x * kTT;
  const uint64_t b0 = (uint64_t)blockIdx.y * kTBy;
  for (unsigned i = threadIdx.x; i < kTT * kTBy; i += kB) {
    const unsigned t = i / kTBy, b = i % kTBy;
    const uint64_t gt = t0 + t, gb = b0 + b;
    tile[t][b] = (gt < nsamps && gb < rowbytes) ? in[gt * rowbytes + gb] : (uint8_t)0;
  }
  __syncthreads();
  for (unsigned i = threadIdx.x; i < kTT * kTBy * K; i += kB) {
    const unsigned t = i % kTT, cl = i / kTT;
    const uint64_t gc = b0 * K + cl, gt = t0 + t;
    if (gc < nchans && gt < nsamps)
      out[gc * nsamps + gt] = (uint8_t)((tile[t][cl / K] >> ((cl % K) * NBITS)) & mask);
  }
}

// delays: ndm x nchans sample delays, -1 for a masked channel.
__global__ void __launch_bounds__(kB) k_dedisp(const uint8_t* __restrict__ rows, uint64_t nsamps_in, uint32_t nchans,
                                               const int32_t* __restrict__ delays, float denom,
                                               float* __restrict__ out, uint64_t out_stride, uint64_t nout) {
  extern __shared__ int32_t d_lds[];
  const uint32_t dm = blockIdx.x;
  for (uint32_t c = threadIdx.x; c < nchans; c += kB) d_lds[c] = delays[(uint64_t)dm * nchans + c];
  __syncthreads();
  const uint64_t ntiles = (nout + kB * kQ - 1) / (kB * kQ);
  for (uint64_t tile = blockIdx.y; tile < ntiles; tile += gridDim.y) {
    const uint64_t t0 = tile * kB * kQ + threadIdx.x;
    uint32_t sum[kQ];
#pragma unroll
    for (unsigned q = 0; q < kQ; ++q) sum[q] = 0u;
    for (uint32_t c = 0; c < nchans; ++c) {
      const int32_t d = d_lds[c];
      if (d < 0) continue;
      const uint8_t* row = rows + (uint64_t)c * nsamps_in + (uint64_t)d;
#pragma unroll
      for (unsigned q = 0; q < kQ; ++q) {
        const uint64_t t = t0 + q * kB;
        if (t < nout) sum[q] += row[t];
      }
    }
#pragma unroll
    for (unsigned q = 0; q < kQ; ++q) {
      const uint64_t t = t0 + q * kB;
      if (t < nout) out[(uint64_t)dm * out_stride + t] = (float)sum[q] / denom;
    }
  }
}

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

int32_t delay_of(float dm, float tab) { return (int32_t)rintf(dm * tab); }

}  // namespace

extern "C" int psx_dedisp_delay_table(uint32_t nchans, float tsamp, float f0, float df, float* table) {
  PSX_REQUIRE(table && nchans > 0 && tsamp > 0 && f0 > 0, "psx_dedisp_delay_table: bad arguments");
  const float b = 1.f / f0;
  for (uint32_t c = 0; c < nchans; ++c) {
    const float a = 1.f / (f0 + (float)c * df);
    table[c] = (float)(4.15e3 / (double)tsamp * (double)(a * a - b * b));
  }
  return psx::ok();
}

extern "C" int psx_dedisp_dm_list(float dm_start, float dm_end, float tsamp, float pulse_width_us, float f0,
                                  float df, uint32_t nchans, float tol, float* out, uint32_t cap, uint32_t* count) {
  PSX_REQUIRE(count && tsamp > 0 && nchans > 0 && tol > 1.0f, "psx_dedisp_dm_list: bad arguments");
  // smearing budget: (dt^2 + width^2)(tol^2 - 1) with the intra-channel and
  // across-band smearing terms at the centre frequency (GHz), times in us
  const double dt = (double)tsamp * 1e6;
  const double ti = (double)pulse_width_us;
  const double f_ghz = ((double)f0 + ((double)(nchans / 2) - 0.5) * (double)df) * 1e-3;
  const double tol2 = (double)tol * (double)tol;
  const double a = 8.3 * (double)df / (f_ghz * f_ghz * f_ghz);
  const double a2 = a * a;
  const double b2 = a2 * (double)((double)nchans * (double)nchans / 16.0);
  const double c = (dt * dt + ti * ti) * (tol2 - 1.0);
  std::vector<float> dms(1, dm_start);
  while (dms.back() < dm_end) {
    const double prev = dms.back();
    const double k = c + tol2 * a2 * prev * prev;
    const double next = (b2 * prev + std::sqrt(-a2 * b2 * prev * prev + (a2 + b2) * k)) / (a2 + b2);
    dms.push_back((float)next);
    PSX_REQUIRE(dms.size() < (1u << 24), "psx_dedisp_dm_list: DM list does not converge");
  }
  *count = (uint32_t)dms.size();
  for (uint32_t i = 0; i < *count && out && i < cap; ++i) out[i] = dms[i];
  if (out && *count > cap) return psx::fail(PSX_ERR_OVERFLOW, "DM list holds %u of %u", cap, *count);
  return psx::ok();
}

extern "C" uint64_t psx_dedisp_max_delay(const float* dm_list, uint32_t ndm, const float* table, uint32_t nchans) {
  if (!dm_list || !table || !ndm || !nchans) return 0;
  const int32_t d = delay_of(dm_list[ndm - 1], table[nchans - 1]);
  return d > 0 ? (uint64_t)d : 0;
}

extern "C" uint64_t psx_dedisp_work_bytes(uint64_t nsamps_in, uint32_t nchans, uint32_t ndm) {
  return round_up(nsamps_in * nchans, 256) + round_up((uint64_t)ndm * nchans * sizeof(int32_t), 256);
}

extern "C" int psx_dedisperse(const uint8_t* d_in, uint64_t nsamps_in, uint32_t nchans, uint32_t nbits,
                              const float* dm_list, uint32_t ndm, const int32_t* killmask, float tsamp, float f0,
                              float df, float* d_out, uint64_t out_stride, void* d_work, psx_stream s) {
  PSX_REQUIRE(d_in && dm_list && d_out && d_work && ndm > 0 && nsamps_in > 0, "psx_dedisperse: bad arguments");
  psx::TraceRange range("Dedisperse");
  PSX_REQUIRE(nchans > 0 && nchans <= kMaxChans, "psx_dedisperse: nchans must be 1..%u", kMaxChans);
  PSX_REQUIRE(nbits == 1 || nbits == 2 || nbits == 4 || nbits == 8, "psx_dedisperse: nbits must be 1, 2, 4 or 8");
  PSX_REQUIRE((uint64_t)nchans * nbits % 8 == 0, "psx_dedisperse: a sample must fill whole bytes");
  PSX_REQUIRE(df < 0, "psx_dedisperse: channels must run from high to low frequency (foff < 0)");
  PSX_REQUIRE(ndm <= 0x7FFFFFFFu, "psx_dedisperse: too many DMs");
  std::vector<float> table(nchans);
  int rc = psx_dedisp_delay_table(nchans, tsamp, f0, df, table.data());
  if (rc) return rc;
  const uint64_t max_delay = psx_dedisp_max_delay(dm_list, ndm, table.data(), nchans);
  PSX_REQUIRE(max_delay < nsamps_in, "psx_dedisperse: %llu samples do not cover the %llu-sample delay",
              (unsigned long long)nsamps_in, (unsigned long long)max_delay);
  const uint64_t nout = nsamps_in - max_delay;
  PSX_REQUIRE(out_stride >= nout, "psx_dedisperse: out_stride below the %llu output samples",
              (unsigned long long)nout);
  std::vector<int32_t> delays((uint64_t)ndm * nchans);
  for (uint32_t d = 0; d < ndm; ++d) {
    PSX_REQUIRE(dm_list[d] >= 0, "psx_dedisperse: negative DM");
    for (uint32_t c = 0; c < nchans; ++c) {
      const int32_t v = delay_of(dm_list[d], table[c]);
      // every read t + delay with t < nout must stay inside the input
      PSX_REQUIRE(v >= 0 && (uint64_t)v <= max_delay,
                  "psx_dedisperse: DM %g needs a delay beyond the list's last DM (sort the list)", dm_list[d]);
      delays[(uint64_t)d * nchans + c] = (killmask && killmask[c] == 0) ? -1 : v;
    }
  }
  uint8_t* rows = static_cast<uint8_t*>(d_work);
  int32_t* d_delays = reinterpret_cast<int32_t*>(rows + round_up(nsamps_in * nchans, 256));
  hipStream_t st = psx::as_stream(s);
  PSX_CHECK_HIP(hipMemcpyAsync(d_delays, delays.data(), delays.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
  const uint64_t rowbytes = (uint64_t)nchans * nbits / 8;
  const uint64_t gx = (nsamps_in + kTT - 1) / kTT, gy = (rowbytes + kTBy - 1) / kTBy;
  PSX_REQUIRE(gx <= 0x7FFFFFFFull && gy <= 65535, "psx_dedisperse: input too large");
  const dim3 ug((unsigned)gx, (unsigned)gy);
  switch (nbits) {
    case 1: k_unpack_t<1><<<ug, kB, 0, st>>>(d_in, nsamps_in, nchans, rows); break;
    case 2: k_unpack_t<2><<<ug, kB, 0, st>>>(d_in, nsamps_in, nchans, rows); break;
    case 4: k_unpack_t<4><<<ug, kB, 0, st>>>(d_in, nsamps_in, nchans, rows); break;
    default: k_unpack_t<8><<<ug, kB, 0, st>>>(d_in, nsamps_in, nchans, rows); break;
  }
  PSX_CHECK_LAUNCH("k_unpack_t");
  const float denom = (float)((1u << nbits) - 1u) * (float)nchans;
  const uint64_t ntiles = (nout + kB * kQ - 1) / (kB * kQ);
  const dim3 dg(ndm, (unsigned)std::min<uint64_t>(ntiles, 65535));
  k_dedisp<<<dg, kB, nchans * sizeof(int32_t), st>>>(rows, nsamps_in, nchans, d_delays, denom, d_out, out_stride, nout);
  PSX_CHECK_LAUNCH("k_dedisp");
  // the delay table's host copy goes out of scope here
  PSX_CHECK_HIP(hipStreamSynchronize(st));
  return psx::ok();
}
