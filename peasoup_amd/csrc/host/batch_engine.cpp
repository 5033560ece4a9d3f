// batch_engine.cpp - host side of the batched acceleration search.
#include "batch_engine.hpp"

#include <algorithm>
#include <chrono>
#include <cstring>
#include <stdexcept>
#include <string>

#include "common.hpp"
#include "kernels_dev.hpp"
#include "peasoup/distiller.hpp"
#include "search_common.hpp"

namespace psx {
namespace batch {

namespace {
constexpr uint32_t kPrefix = 8192;  // candidates copied back with every batch

void hip_ok(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
void psx_ok(int rc, const char* what) {
  if (rc != PSX_OK) throw std::runtime_error(std::string(what) + ": " + psx_last_error_string());
}
template <class T>
void dalloc(T** p, uint64_t n) {
  hip_ok(hipMalloc((void**)p, sizeof(T) * (n ? n : 1)), "hipMalloc");
}
template <class T>
void halloc(T** p, uint64_t n) {
  hip_ok(hipHostMalloc((void**)p, sizeof(T) * (n ? n : 1), hipHostMallocDefault), "hipHostMalloc");
}
int64_t round_up(int64_t v, int64_t a) { return (v + a - 1) / a * a; }
}  // namespace

uint64_t Engine::device_bytes(uint64_t N, uint32_t B, uint32_t cap) {
  const uint64_t M = N / 2, nb = M + 1;
  const uint64_t pst = (uint64_t)round_up((int64_t)nb, 64);
  uint64_t b = 0;
  b += 2 * M * B * sizeof(float2);      // FFT work + Z
  b += (pst * B + kHarmWaveBins) * sizeof(float);  // P
  b += (uint64_t)bound_stride((int64_t)pst) * B * sizeof(uint16_t);  // Q
  b += 2 * (uint64_t)cap * sizeof(Cand16);
  b += nb * (sizeof(float2) + 2 * sizeof(float)) + nb / 4 * sizeof(float);  // prologue X, A, median
  b += 2 * M * sizeof(float2) + 16 * 1024 * 1024;                          // prologue plans + tables
  return b;
}

Engine::Engine(const psx_search_args& a, const EngineOptions& opt, hipStream_t stream)
    : args(a), N(a.size), M(a.size / 2), nb(a.size / 2 + 1), B(opt.batch), cap(opt.cand_capacity),
      F(a.nharmonics), use_graph(opt.use_graph), s(stream) {
  if (B < 1 || B > 1024) throw std::runtime_error("batch must be 1..1024");
  if (cap < kPrefix) cap = kPrefix;
  pstride = round_up((int64_t)nb, 64);
  psx_ok(psx::fft::plan_create_with(N, PSX_FFT_R2C, B, opt.decomp, &plan), "batched R2C plan");
  rt = psx::fft::root_table(plan);
  if (opt.spectrum_layout < PSX_SPECTRUM_AUTO || opt.spectrum_layout > PSX_SPECTRUM_ROWS)
    throw std::runtime_error("spectrum_layout must be PSX_SPECTRUM_AUTO, _NATURAL or _ROWS");
  const bool rows_ok =
      psx::fft::rows_layout(plan, &rows_R, &rows_L) && rows_L % 64 == 0 && rows_R % 2 == 0 && rows_R >= 4;
  if (opt.spectrum_layout == PSX_SPECTRUM_ROWS && !rows_ok)
    throw std::runtime_error("spectrum_layout ROWS: this FFT plan has no row-layout last pass");
  rows_out = rows_ok && opt.spectrum_layout != PSX_SPECTRUM_NATURAL;
  if (opt.harmonic_pass != PSX_HARMONIC_BOUND && opt.harmonic_pass != PSX_HARMONIC_EXACT)
    throw std::runtime_error("harmonic_pass must be PSX_HARMONIC_BOUND or PSX_HARMONIC_EXACT");
  bound = opt.harmonic_pass == PSX_HARMONIC_BOUND && F > 0;
  // the row-layout pass pairs bins (2u, 2u+1) by rows, which needs R % 4 == 0
  // (true of every plan with that layout: R is 2^k or 3*2^k, k >= 5)
  fused = bound && rows_out && rows_R % 4 == 0;
  dalloc(&Z, M * B);
  if (!fused) dalloc(&P, (uint64_t)pstride * B + kHarmWaveBins);
  if (bound) {
    // the pooling pass writes each row's bounds up to bin nb; the slack past
    // them (read only for bins past the window) holds -inf
    qstride = bound_stride(pstride);
    dalloc(&Q, (uint64_t)qstride * B);
    hip_ok(hipMemsetD16(Q, 0xFF80u, (size_t)qstride * B), "bound spectrum fill");  // bfloat16 -inf
    if (fused) {
      dalloc((uint8_t**)&rows_src, kRowsSourceBytes);
      psx_ok(write_rows_source(RowsSource{Z, rows_R, rows_L, rt}, rows_src), "rows source");
    }
  }
  for (Slot& sl : slots) {
    dalloc(&sl.d_desc, B);
    halloc(&sl.h_desc, B);
    dalloc(&sl.d_cands, cap);
    halloc(&sl.h_cands, kPrefix);
    dalloc(&sl.d_count, 1);
    halloc(&sl.h_count, 1);
    hip_ok(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "hipEventCreate");
  }
  // PeakFinder windows and bin->frequency factors per harmonic number
  // (pipeline_multi.cu:130-131: float tobs = size*tsamp; bin_width = 1.0/tobs)
  const float tobs = (unsigned int)N * args.tsamp;
  const float bin_width = 1.0 / tobs;
  win.lo = INT64_MAX;
  win.hi = 0;
  for (uint32_t nh = 0; nh <= F; ++nh) {
    const peasoup::PeakWindow w = peasoup::peak_window((unsigned)nb, (double)bin_width, nh, args.min_freq,
                                                       args.max_freq);
    win.start[nh] = std::max(w.start, 0);
    win.end[nh] = std::min(w.end, (int)nb);
    factor[nh] = w.factor;
    if (win.end[nh] > win.start[nh]) {
      win.lo = std::min<int64_t>(win.lo, win.start[nh]);
      win.hi = std::max<int64_t>(win.hi, win.end[nh]);
    }
  }
  if (win.hi <= win.lo) win.lo = win.hi = 0;
  win.thr = args.min_snr;
  // prologue resources (plain path)
  psx_ok(psx_fft_plan_create(N, PSX_FFT_R2C, 1, &p_r2c), "prologue R2C plan");
  psx_ok(psx_fft_plan_create(N, PSX_FFT_C2R, 1, &p_c2r), "prologue C2R plan");
  dalloc(&X, nb);
  dalloc(&A, nb);
  dalloc(&med, nb);
  dalloc(&med_work, psx_deredden_work_floats(nb));
  dalloc((uint8_t**)&red_work, psx_reduce_work_bytes());
  dalloc(&d_base, 1);
  dalloc(&d_sums, 1);
}

Engine::~Engine() {
  if (s) (void)hipStreamSynchronize(s);
  for (Slot& sl : slots) {
    for (auto& g : sl.graphs) (void)hipGraphExecDestroy(g.second);
    for (auto& g : sl.defs) (void)hipGraphDestroy(g.second);
    (void)hipFree(sl.d_desc);
    (void)hipHostFree(sl.h_desc);
    (void)hipFree(sl.d_cands);
    (void)hipHostFree(sl.h_cands);
    (void)hipFree(sl.d_count);
    (void)hipHostFree(sl.h_count);
    if (sl.done) (void)hipEventDestroy(sl.done);
  }
  (void)hipFree(Z);
  (void)hipFree(P);
  (void)hipFree(Q);
  (void)hipFree(rows_src);
  psx_fft_plan_destroy(plan);
  psx_fft_plan_destroy(p_r2c);
  psx_fft_plan_destroy(p_c2r);
  (void)hipFree(X);
  (void)hipFree(A);
  (void)hipFree(med);
  (void)hipFree(med_work);
  (void)hipFree(red_work);
  (void)hipFree(d_base);
  (void)hipFree(d_sums);
}

void Engine::prologue(const float* d_series, uint64_t n_real, float* d_tim, float* mean, float* std,
                      const Zapper* zap) {
  // pipeline_multi.cu:170-238
  const uint64_t ncopy = std::min<uint64_t>(n_real, N);
  if (d_series != d_tim) hip_ok(hipMemcpyAsync(d_tim, d_series, ncopy * sizeof(float), hipMemcpyDeviceToDevice, s),
                                "prologue copy");
  psx_ok(psx_remove_baseline(d_tim, ncopy, red_work, s), "remove_baseline");
  if (N > n_real) psx_ok(psx_fill(d_tim, n_real, N, 0.0f, s), "fill");
  const float tobs = (unsigned int)N * args.tsamp;
  const float bin_width = 1.0 / tobs;
  psx_ok(psx_fft_exec(p_r2c, d_tim, X, s), "prologue R2C");
  psx_ok(psx_form_amplitude(X, A, nb, s), "form");
  psx_ok(psx_deredden_median(A, nb, (double)bin_width, 0.05f, 0.5f, med, med_work, s), "calculate_median");
  psx_ok(psx_divide_c_by_f(X, med, nb, s), "deredden");
  if (zap) zap->zap_on(X, bin_width, nb, s);
  psx_ok(psx_form_interbin(X, A, nb, s), "form_interpolated");
  float m = 0, r = 0;
  psx_ok(psx_mean_rms(A, nb, 0, &m, &r, red_work, s), "stats");
  *mean = m;
  *std = std::sqrt(r * r - m * m);
  psx_ok(psx_fft_exec(p_c2r, X, d_tim, s), "prologue C2R");
}

// Spectrum pass: from the natural half-length spectra, or from the
// transposed rows when the plan's last pass writes those (its 32/64-byte
// column runs cost more than the spectrum pass's tiled transpose).
// In the bound-filtered form the row-layout pass also writes the bound
// spectrum and the fundamental's candidates; after the natural-layout pass a
// pooling pass computes them from P.
void Engine::spectrum(Slot& sl, uint32_t n) {
  if (rows_out) {
    const SpecBound sb{Q, qstride, win, sl.d_cands, cap, sl.d_count};
    psx_ok(launch_spec_rows(Z, fused ? nullptr : P, rows_R, rows_L, pstride, rt, sl.d_desc, n, s,
                            fused ? &sb : nullptr),
           "spec_rows");
  } else {
    psx_ok(launch_spec_post(Z, P, M, pstride, rt, sl.d_desc, n, s), "spec_post");
  }
  if (bound && !fused)
    psx_ok(launch_pool_peaks(P, pstride, nb, win, Q, qstride, sl.d_desc, n, sl.d_cands, cap, sl.d_count, s),
           "pool_peaks");
}

namespace {
// bfloat16 bits of the smallest bfloat16 >= x (x not NaN), computed as the
// nearest-up rounding of the float's top half: exact for every float
uint16_t bf16_up_bits(float x) {
  uint32_t u;
  std::memcpy(&u, &x, 4);
  if (u & 0x80000000u) return (uint16_t)(u >> 16);  // negative: toward zero is up
  const float t = [&] {
    const uint32_t top = u & 0xFFFF0000u;
    float f;
    std::memcpy(&f, &top, 4);
    return f;
  }();
  if (t == x) return (uint16_t)(u >> 16);
  return (uint16_t)((u >> 16) + 1u);  // the next bfloat16 up (+inf past the largest finite)
}
}  // namespace

void Engine::audit_bounds(const TrialWork* trials, uint32_t n, BoundAudit& out) {
  if (!bound) throw std::runtime_error("audit_bounds: the engine runs the exact harmonic pass");
  if (n < 1 || n > B) throw std::runtime_error("audit_bounds: 1..B trials");
  if (!pending.empty() || !ready.empty()) throw std::runtime_error("audit_bounds: engine busy");
  Slot& sl = slots[0];
  fill_slot(sl, trials, n);
  // the fused form writes no spectra: the audit has the exact-form spectrum
  // pass write them from the same FFT output after the bound-form pass
  if (!P) dalloc(&P, (uint64_t)pstride * B + kHarmWaveBins);
  hip_ok(hipMemcpyAsync(sl.d_desc, sl.h_desc, n * sizeof(psx::fft::TrialDesc), hipMemcpyHostToDevice, s), "desc");
  psx_ok(launch_zero(sl.d_count, s), "zero");
  psx_ok(psx::fft::exec_fused_r2c(plan, sl.d_desc, n, Z, s, rows_out), "fused FFT");
  spectrum(sl, n);
  if (fused) psx_ok(launch_spec_rows(Z, P, rows_R, rows_L, pstride, rt, sl.d_desc, n, s, nullptr), "spec_rows");
  std::vector<float> hp((size_t)pstride * n);
  std::vector<uint16_t> hq((size_t)qstride * n);
  uint32_t count = 0;
  hip_ok(hipMemcpyAsync(hp.data(), P, hp.size() * sizeof(float), hipMemcpyDeviceToHost, s), "P copy");
  hip_ok(hipMemcpyAsync(hq.data(), Q, hq.size() * sizeof(uint16_t), hipMemcpyDeviceToHost, s), "Q copy");
  hip_ok(hipMemcpyAsync(&count, sl.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "count copy");
  hip_ok(hipStreamSynchronize(s), "audit sync");
  out = BoundAudit{};
  const uint64_t nq = (nb + 1) / 2;
  for (uint32_t t = 0; t < n; ++t) {
    const float* pr = hp.data() + (size_t)t * pstride;
    const uint16_t* qr = hq.data() + (size_t)t * qstride;
    for (uint64_t u = 0; u < nq; ++u) {
      const float a = pr[2 * u];
      const float m = 2 * u + 1 < nb ? std::max(a, pr[2 * u + 1]) : a;
      out.entries++;
      // a valid bound: its float value is at least the pair's maximum; the
      // tight one is bf16_up_bits(m), and the fused pass (which bounds each
      // value from approximate arithmetic) may be one step above it
      const uint16_t tight = bf16_up_bits(m);
      const uint32_t bits = (uint32_t)qr[u] << 16;
      float qv;
      std::memcpy(&qv, &bits, 4);
      if (!(qv >= m)) out.mismatches++;
      else if (qr[u] != tight) out.loose++;
    }
    for (uint64_t k = (uint64_t)std::max(win.start[0], 0); k < (uint64_t)std::max(win.end[0], 0); ++k)
      if (pr[k] > win.thr) out.fundamentals++;
    // the harmonic pass's bound sums, in its order and float arithmetic
    float ft[6];
    for (uint32_t k = 1; k <= F; ++k) ft[k] = fold_threshold_of((int)k, win.thr);
    auto qv = [&](uint64_t x) {
      const uint32_t bits = (uint32_t)qr[x / 2] << 16;
      float f;
      std::memcpy(&f, &bits, 4);
      return f;
    };
    for (uint64_t i = (uint64_t)win.lo; i < (uint64_t)win.hi; ++i) {
      float val = qv(i);
      bool cross = false;
      int h = 0;
      for (uint32_t k = 1; k <= F; ++k) {
        const uint64_t c = 1ull << (k - 1);
        // fold k's odd multiples j/2^k, fold 2 taking 3/4 before 1/4
        for (uint64_t jj = 0; jj < c; ++jj, ++h) {
          const uint64_t j = (k == 2) ? (jj == 0 ? 3 : 1) : 2 * jj + 1;
          val += qv((j * i + c) >> k);
        }
        cross |= val > ft[k];
      }
      out.window++;
      if (cross) out.crossing++;
    }
  }
  out.appended = count;
}


// Harmonic pass: the bound-filtered form (the folds from the bounds the
// spectrum stage wrote) or the exact gathers of every addend.
void Engine::harmonics(Slot& sl, uint32_t n) {
  if (bound) {
    const RowsSource rs{Z, rows_R, rows_L, rt};
    psx_ok(launch_harm_bound(fused ? nullptr : P, pstride, Q, qstride, win, F, sl.d_desc, n, sl.d_cands, cap,
                             sl.d_count, s, fused ? &rs : nullptr, rows_src),
           "harm_bound");
  } else {
    psx_ok(launch_harm_peaks(P, pstride, win, F, sl.d_desc, n, sl.d_cands, cap, sl.d_count, s), "harm_peaks");
  }
}

void Engine::prologue_async(const float* d_series, uint64_t n_real, float* d_tim, PrologueResult* res,
                            hipStream_t ps, const Zapper* zap) {
  // prologue() step for step, on ps
  const uint64_t ncopy = std::min<uint64_t>(n_real, N);
  if (d_series != d_tim)
    hip_ok(hipMemcpyAsync(d_tim, d_series, ncopy * sizeof(float), hipMemcpyDeviceToDevice, ps), "prologue copy");
  psx_ok(psx::remove_baseline_async(d_tim, ncopy, red_work, d_base, kBaselineSteps, ps), "remove_baseline");
  if (N > n_real) psx_ok(psx_fill(d_tim, n_real, N, 0.0f, ps), "fill");
  const float tobs = (unsigned int)N * args.tsamp;
  const float bin_width = 1.0 / tobs;
  psx_ok(psx_fft_exec(p_r2c, d_tim, X, ps), "prologue R2C");
  psx_ok(psx_form_amplitude(X, A, nb, ps), "form");
  psx_ok(psx_deredden_median(A, nb, (double)bin_width, 0.05f, 0.5f, med, med_work, ps), "calculate_median");
  psx_ok(psx_divide_c_by_f(X, med, nb, ps), "deredden");
  if (zap) zap->zap_on(X, bin_width, nb, ps);
  psx_ok(psx_form_interbin(X, A, nb, ps), "form_interpolated");
  psx_ok(psx::sum_sq_async(A, 0, nb, red_work, d_sums, ps), "stats");
  hip_ok(hipMemcpyAsync(&res->base, d_base, sizeof(BaselineState), hipMemcpyDeviceToHost, ps), "baseline state");
  hip_ok(hipMemcpyAsync(&res->sums, d_sums, sizeof(double2), hipMemcpyDeviceToHost, ps), "stats copy");
  psx_ok(psx_fft_exec(p_c2r, X, d_tim, ps), "prologue C2R");
}

// psx_mean_rms's arithmetic on the sums, then prologue()'s std.
void Engine::stats_of(const PrologueResult& res, float* mean, float* std) const {
  const float cnt = (float)(int64_t)nb;
  const float m = (float)res.sums.x / cnt;
  const float r = sqrtf((float)res.sums.y / cnt);
  *mean = m;
  *std = std::sqrt(r * r - m * m);
}

void Engine::enqueue(Slot& sl, uint32_t n) { launch(sl, n, true); }

void Engine::prepare(uint32_t n) {
  if (!use_graph || n == 0 || n > B) return;
  for (Slot& sl : slots) launch(sl, n, false);
}

void Engine::launch(Slot& sl, uint32_t n, bool run) {
  auto record = [&]() {
    hip_ok(hipMemcpyAsync(sl.d_desc, sl.h_desc, n * sizeof(psx::fft::TrialDesc), hipMemcpyHostToDevice, s),
           "descriptor upload");
    psx_ok(launch_zero(sl.d_count, s), "zero");
    psx_ok(psx::fft::exec_fused_r2c(plan, sl.d_desc, n, Z, s, rows_out), "fused FFT");
    spectrum(sl, n);
    harmonics(sl, n);
    hip_ok(hipMemcpyAsync(sl.h_count, sl.d_count, sizeof(uint32_t), hipMemcpyDeviceToHost, s), "count copy");
    hip_ok(hipMemcpyAsync(sl.h_cands, sl.d_cands, kPrefix * sizeof(Cand16), hipMemcpyDeviceToHost, s),
           "candidate copy");
  };
  if (use_graph) {
    auto it = sl.graphs.find(n);
    if (it == sl.graphs.end()) {
      hipGraph_t g;
      hip_ok(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal), "begin capture");
      try {
        record();
      } catch (...) {
        (void)hipStreamEndCapture(s, &g);
        throw;
      }
      hip_ok(hipStreamEndCapture(s, &g), "end capture");
      hipGraphExec_t ge;
      const hipError_t e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
      if (e != hipSuccess) (void)hipGraphDestroy(g);
      hip_ok(e, "graph instantiate");
      sl.defs.emplace(n, g);
      it = sl.graphs.emplace(n, ge).first;
    }
    if (!run) return;
    hip_ok(hipGraphLaunch(it->second, s), "graph launch");
  } else {
    record();
  }
  hip_ok(hipEventRecord(sl.done, s), "event record");
}

void Engine::submit(const TrialWork* trials, uint32_t n) {
  if (n == 0) return;
  if (n > B) throw std::runtime_error("submit: more trials than the batch size");
  if (pending.size() == 2) {
    ready.emplace_back();
    collect_oldest(ready.back());
  }
  Slot& sl = slots[next_slot];
  fill_slot(sl, trials, n);
  enqueue(sl, n);
  pending.push_back(next_slot);
  next_slot ^= 1;
}

int Engine::stage(const TrialWork* trials, uint32_t n) {
  if (!use_graph) throw std::runtime_error("stage: mixed dispatch needs launch graphs");
  if (n == 0 || n > B) throw std::runtime_error("stage: 1..B trials");
  if (pending.size() == 2) {
    ready.emplace_back();
    collect_oldest(ready.back());
  }
  Slot& sl = slots[next_slot];
  fill_slot(sl, trials, n);
  launch(sl, n, false);
  return next_slot;
}

hipGraph_t Engine::graph_def(int slot, uint32_t n) const {
  const auto it = slots[slot].defs.find(n);
  if (it == slots[slot].defs.end()) throw std::runtime_error("graph_def: no graph for this batch");
  return it->second;
}

void Engine::launched(int slot, hipStream_t ls) {
  if (slot != next_slot) throw std::runtime_error("launched: slot was not the staged one");
  hip_ok(hipEventRecord(slots[slot].done, ls), "event record");
  pending.push_back(slot);
  next_slot ^= 1;
}

// Trial descriptors of a batch into the slot's pinned buffer (the batch's
// graph uploads them).
void Engine::fill_slot(Slot& sl, const TrialWork* trials, uint32_t n) {
  sl.meta.assign(trials, trials + n);
  for (uint32_t t = 0; t < n; ++t) {
    const TrialWork& w = trials[t];
    psx::fft::TrialDesc d;
    d.tim = w.tim;
    d.af = psx::accel_fact(w.acc, args.tsamp);
    if (!(fabs(d.af) * (double)N < 1.0)) throw std::runtime_error("acceleration too large for the series length");
    d.mean_n = w.mean * (unsigned int)N;  // stats::normalise(pspec, mean*size, std*size, ...)
    d.std_n = w.std * (unsigned int)N;
    d.slot = t;
    d.pad = 0;
    sl.h_desc[t] = d;
  }
}

bool Engine::collect(std::vector<std::vector<Candidate>>& per_trial) {
  if (!ready.empty()) {
    per_trial = std::move(ready.front());
    ready.erase(ready.begin());
    return true;
  }
  return collect_oldest(per_trial);
}

bool Engine::ready_or_done() {
  if (!ready.empty()) return true;
  if (pending.empty()) return false;
  const hipError_t q = hipEventQuery(slots[pending.front()].done);
  if (q == hipErrorNotReady) return false;
  hip_ok(q, "batch query");
  return true;
}

bool Engine::collect_oldest(std::vector<std::vector<Candidate>>& per_trial) {
  if (pending.empty()) return false;
  Slot& sl = slots[pending.front()];
  pending.erase(pending.begin());
  hip_ok(hipEventSynchronize(sl.done), "batch wait");
  process(sl, per_trial);
  return true;
}

void Engine::process(Slot& sl, std::vector<std::vector<Candidate>>& per_trial) {
  const uint32_t n = (uint32_t)sl.meta.size();
  const uint32_t count = *sl.h_count;
  if (count > cap)
    throw std::runtime_error("batched search: " + std::to_string(count) + " bins above threshold exceed the " +
                             std::to_string(cap) + "-entry candidate buffer");
  std::vector<Cand16> all(sl.h_cands, sl.h_cands + std::min(count, kPrefix));
  if (count > kPrefix) {
    all.resize(count);
    hip_ok(hipMemcpy(all.data() + kPrefix, sl.d_cands + kPrefix, (count - kPrefix) * sizeof(Cand16),
                     hipMemcpyDeviceToHost),
           "candidate tail copy");
  }
  // order by (trial, harmonic, bin): the stable order device_find_peaks
  // produces for each spectrum
  std::sort(all.begin(), all.end(), [](const Cand16& a, const Cand16& b) {
    if (a.slot != b.slot) return a.slot < b.slot;
    if (a.nh != b.nh) return a.nh < b.nh;
    return a.idx < b.idx;
  });
  per_trial.assign(n, {});
  HarmonicDistiller harm(args.freq_tol, args.max_harm, false);
  std::vector<int> idxs, pidx;
  std::vector<float> snrs, psnr, pfreq;
  size_t pos = 0;
  for (uint32_t t = 0; t < n; ++t) {
    const TrialWork& w = sl.meta[t];
    SpectrumCandidates sc(w.dm, w.dm_idx, w.acc);
    for (uint32_t nh = 0; nh <= F; ++nh) {
      idxs.clear();
      snrs.clear();
      while (pos < all.size() && all[pos].slot == t && all[pos].nh == nh) {
        idxs.push_back(all[pos].idx);
        snrs.push_back(all[pos].snr);
        ++pos;
      }
      // PeakFinder host buffers hold 100000 entries (peakfinder.hpp:61-67)
      if (idxs.size() > 100000) throw std::runtime_error("more than 100000 bins above threshold in one spectrum");
      pidx.resize(idxs.size() + 1);
      psnr.resize(idxs.size() + 1);
      const int np = peasoup::identify_unique_peaks(idxs.data(), snrs.data(), (int)idxs.size(), args.min_gap,
                                                    pidx.data(), psnr.data());
      pfreq.resize(np + 1);
      for (int i = 0; i < np; ++i) pfreq[i] = pidx[i] * factor[nh];
      sc.append(psnr.data(), pfreq.data(), (int)nh, np);
    }
    per_trial[t] = harm.distill(sc.cands);
  }
}

void Engine::profile(const TrialWork* trials, uint32_t n, int reps, StageTimes& out) {
  if (n < 1 || n > B) throw std::runtime_error("profile: 1..B trials");
  if (!pending.empty() || !ready.empty()) throw std::runtime_error("profile: engine busy");
  Slot& sl = slots[0];
  for (uint32_t t = 0; t < n; ++t) {
    psx::fft::TrialDesc d;
    d.tim = trials[t].tim;
    d.af = psx::accel_fact(trials[t].acc, args.tsamp);
    d.mean_n = trials[t].mean * (unsigned int)N;
    d.std_n = trials[t].std * (unsigned int)N;
    d.slot = t;
    d.pad = 0;
    sl.h_desc[t] = d;
  }
  hip_ok(hipMemcpyAsync(sl.d_desc, sl.h_desc, n * sizeof(psx::fft::TrialDesc), hipMemcpyHostToDevice, s), "desc");
  const int np = psx::fft::num_passes(plan);
  const int nstage = np + 2;
  std::vector<hipEvent_t> ev(nstage + 1);
  for (auto& e : ev) hip_ok(hipEventCreate(&e), "event");
  std::vector<double> acc(nstage, 0.0);
  // Each stage is timed over kRun back-to-back launches of itself between
  // two events, in pipeline order (every launch recomputes the same output
  // from the same input), so the event markers' launch gaps are spread over
  // kRun launches instead of landing on every stage of a single batch.  A
  // pass that works in place is launched once: a second launch would
  // transform its own output, and every later stage would time garbage.
  constexpr int kRun = 4;
  auto runs = [&](int k) { return k < np && psx::fft::pass_in_place(plan, k) ? 1 : kRun; };
  auto stage = [&](int k) {
    if (k < np) {
      psx_ok(psx::fft::exec_fused_pass(plan, k, sl.d_desc, n, Z, s, rows_out), "fused pass");
    } else if (k == np) {
      spectrum(sl, n);
    } else {
      harmonics(sl, n);
    }
  };
  for (int r = 0; r < reps + 1; ++r) {
    psx_ok(launch_zero(sl.d_count, s), "zero");
    hip_ok(hipEventRecord(ev[0], s), "event");
    for (int k = 0; k < nstage; ++k) {
      for (int i = 0; i < (r == 0 ? 1 : runs(k)); ++i) stage(k);
      hip_ok(hipEventRecord(ev[k + 1], s), "event");
    }
    hip_ok(hipEventSynchronize(ev[nstage]), "event sync");
    if (r == 0) continue;  // warm-up repetition (one launch per stage)
    for (int k = 0; k < nstage; ++k) {
      float ms = 0;
      hip_ok(hipEventElapsedTime(&ms, ev[k], ev[k + 1]), "elapsed");
      acc[k] += ms / runs(k);
    }
  }
  for (auto& e : ev) (void)hipEventDestroy(e);
  out.ms.assign(nstage, 0.0);
  for (int k = 0; k < nstage; ++k) out.ms[k] = acc[k] / reps;
  out.npass = np;
  const psx::fft::Decomp& d = psx::fft::plan_decomp(plan);
  out.desc.clear();
  for (int q = 0; q < d.npass; ++q)
    out.desc += (q ? "x" : "") + std::to_string(d.L[q]) + "/C" + std::to_string(d.C[q]);
  if (rows_out) out.desc += " rows";
}

}  // namespace batch
}  // namespace psx
