// batch_search.cpp - whole-search driver over the batched engine: the
// Worker's DM loop (src/pipeline_multi.cu:155-295) restructured so that the
// DM prologues run ahead on a stream of their own and the acceleration
// trials of every DM stream through fused, graph-launched batches as soon as
// their DM's prologue is done.  Candidate semantics are the
// Worker's: per-trial HarmonicDistiller, per-DM AccelerationDistiller over
// the trials in acceleration-list order.
#include <chrono>
#include <cstring>
#include <exception>
#include <functional>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "batch_engine.hpp"
#include "common.hpp"
#include "peasoup/distiller.hpp"
#include "search_common.hpp"
#include "trace.hpp"

using psx::batch::Engine;
using psx::batch::EngineOptions;
using psx::batch::TrialWork;

namespace {
using clk = std::chrono::steady_clock;
double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }
void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}
}  // namespace

namespace psx {
namespace search {

// Runs the accel trials of the given DMs (device series already prologued
// into tims) and returns the accel-distilled candidates per DM.
void run_trials(Engine& eng, const psx_search_args& a, const std::vector<const float*>& tims,
                const std::vector<float>& means, const std::vector<float>& stds, const std::vector<float>& dms,
                const std::vector<int>& dm_idx, const std::vector<std::vector<float>>& accs,
                std::vector<std::vector<Candidate>>& per_dm, const std::function<void(size_t)>& ready) {
  const size_t ndm = tims.size();
  std::vector<TrialWork> all;
  std::vector<std::pair<size_t, size_t>> where;  // (dm, acc index) of each trial
  for (size_t d = 0; d < ndm; ++d)
    for (size_t j = 0; j < accs[d].size(); ++j) {
      all.push_back(TrialWork{tims[d], accs[d][j], means[d], stds[d], dms[d], dm_idx[d]});
      where.emplace_back(d, j);
    }
  std::vector<std::vector<std::vector<Candidate>>> per_trial(ndm);
  for (size_t d = 0; d < ndm; ++d) per_trial[d].resize(accs[d].size());
  const uint32_t B = eng.batch();
  size_t next = 0, done = 0;
  std::vector<size_t> starts;  // first trial of each submitted batch
  std::vector<std::vector<Candidate>> got;
  std::vector<char> is_ready(ndm, ready ? 0 : 1);
  while (done < all.size()) {
    while (next < all.size() && eng.in_flight() < 2) {
      const uint32_t n = (uint32_t)std::min<size_t>(B, all.size() - next);
      // a DM's statistics (and its series on the device) exist once its
      // prologue has finished: ready(d) waits for that
      for (size_t t = next; t < next + n; ++t) {
        const size_t d = where[t].first;
        if (!is_ready[d]) {
          ready(d);
          is_ready[d] = 1;
        }
        all[t].mean = means[d];
        all[t].std = stds[d];
      }
      eng.submit(&all[next], n);
      starts.push_back(next);
      next += n;
    }
    if (!eng.collect(got)) break;
    const size_t first = starts.front();
    starts.erase(starts.begin());
    for (size_t t = 0; t < got.size(); ++t) {
      const auto& w = where[first + t];
      per_trial[w.first][w.second] = std::move(got[t]);
    }
    done += got.size();
  }
  const float tobs = (unsigned int)a.size * a.tsamp;
  per_dm.assign(ndm, {});
  for (size_t d = 0; d < ndm; ++d) {
    std::vector<Candidate> accel_trial_cands;
    for (auto& v : per_trial[d]) accel_trial_cands.insert(accel_trial_cands.end(), v.begin(), v.end());
    AccelerationDistiller acc_still(tobs, a.freq_tol, true);
    per_dm[d] = acc_still.distill(accel_trial_cands);
  }
}

}  // namespace search
}  // namespace psx

namespace {
bool engine_options(const psx_batch_opts* opts, uint64_t n, EngineOptions& eo, psx::fft::Decomp& forced) {
  eo.batch = psx::batch::default_batch(n);
  if (opts) {
    eo.batch = opts->batch ? opts->batch : eo.batch;
    eo.cand_capacity = opts->cand_capacity ? opts->cand_capacity : eo.cand_capacity;
    eo.use_graph = opts->use_graph != 0;
    eo.spectrum_layout = opts->spectrum_layout;
    eo.harmonic_pass = opts->harmonic_pass;
    if (opts->fft_npass > 0) {
      forced.npass = opts->fft_npass;
      for (int q = 0; q < 3; ++q) {
        forced.L[q] = opts->fft_len[q];
        forced.C[q] = opts->fft_cols[q];
      }
      eo.decomp = &forced;
    }
  }
  return true;
}

// The engine (with its stream) that the last call asking for it left
// behind: its device buffers, FFT plans and launch graphs serve the next
// call with the same device, search arguments and options.  A call takes it
// out of the cache while it runs, so concurrent calls never share one.
struct EngineKey {
  int device = -1;
  psx_search_args args;
  uint32_t batch = 0, cand_capacity = 0;
  bool use_graph = false, forced = false;
  int spectrum_layout = 0;
  int harmonic_pass = 0;
  psx::fft::Decomp decomp;
};
EngineKey engine_key(const psx_search_args& a, const EngineOptions& eo) {
  EngineKey k;
  std::memset(&k.args, 0, sizeof(k.args));
  k.args = a;
  (void)hipGetDevice(&k.device);
  k.batch = eo.batch;
  k.cand_capacity = eo.cand_capacity;
  k.use_graph = eo.use_graph;
  k.spectrum_layout = eo.spectrum_layout;
  k.harmonic_pass = eo.harmonic_pass;
  k.forced = eo.decomp != nullptr;
  if (eo.decomp) k.decomp = *eo.decomp;
  return k;
}
bool same_key(const EngineKey& x, const EngineKey& y) {
  return x.device == y.device && std::memcmp(&x.args, &y.args, sizeof(x.args)) == 0 && x.batch == y.batch &&
         x.cand_capacity == y.cand_capacity && x.use_graph == y.use_graph && x.forced == y.forced &&
         x.spectrum_layout == y.spectrum_layout && x.harmonic_pass == y.harmonic_pass &&
         (!x.forced || (x.decomp.npass == y.decomp.npass &&
                        std::memcmp(x.decomp.L, y.decomp.L, sizeof(x.decomp.L)) == 0 &&
                        std::memcmp(x.decomp.C, y.decomp.C, sizeof(x.decomp.C)) == 0));
}
struct KeptEngine {
  std::unique_ptr<Engine> eng;
  hipStream_t s = nullptr;
  EngineKey key;
  void reset() {
    eng.reset();
    if (s) (void)hipStreamDestroy(s);
    s = nullptr;
  }
};
// never destroyed: an engine still kept at exit is left to the process
// teardown (freeing it from a static destructor could run after the HIP
// runtime's own teardown)
std::mutex& g_kept_mu = *new std::mutex;
KeptEngine& g_kept = *new KeptEngine;

// Takes the kept engine when its key matches (else leaves the cache alone).
bool take_kept(const EngineKey& k, std::unique_ptr<Engine>& eng, hipStream_t& s) {
  std::lock_guard<std::mutex> lock(g_kept_mu);
  if (!g_kept.eng || !same_key(g_kept.key, k)) return false;
  eng = std::move(g_kept.eng);
  s = g_kept.s;
  g_kept.s = nullptr;
  return true;
}
// Keeps an idle engine; the one it replaces is freed outside the lock.
void keep(std::unique_ptr<Engine> eng, hipStream_t s, const EngineKey& k) {
  KeptEngine old;
  {
    std::lock_guard<std::mutex> lock(g_kept_mu);
    old.eng = std::move(g_kept.eng);
    old.s = g_kept.s;
    g_kept.eng = std::move(eng);
    g_kept.s = s;
    g_kept.key = k;
  }
  old.reset();
}
}  // namespace

extern "C" int psx_batch_release(void) {
  KeptEngine old;
  {
    std::lock_guard<std::mutex> lock(g_kept_mu);
    old.eng = std::move(g_kept.eng);
    old.s = g_kept.s;
    g_kept.s = nullptr;
  }
  old.reset();
  return psx::ok();
}

extern "C" int psx_batch_profile(const psx_search_args* args, const psx_batch_opts* opts, const float* d_series,
                                 uint64_t nsamps, const float* accs, uint32_t naccs, int reps, double* stage_ms,
                                 int* nstage, char* desc, size_t cap) {
  int rc = psx::search::validate_args(args);
  if (rc) return rc;
  PSX_REQUIRE(d_series && accs && naccs > 0 && reps > 0 && stage_ms && nstage, "psx_batch_profile: bad arguments");
  hipStream_t s = nullptr;
  float* d_tim = nullptr;
  try {
    EngineOptions eo;
    psx::fft::Decomp forced;
    engine_options(opts, args->size, eo, forced);
    PSX_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<Engine> eng(new Engine(*args, eo, s));
    PSX_CHECK_HIP(hipMalloc(&d_tim, sizeof(float) * args->size));
    std::unique_ptr<Zapper> zap;
    if (args->nbirdies) zap.reset(new Zapper(args->birdie_freq, args->birdie_width, args->nbirdies, s));
    float mean, sd;
    eng->prologue(d_series, nsamps, d_tim, &mean, &sd, zap.get());
    const uint32_t n = std::min(naccs, eng->batch());
    std::vector<TrialWork> tw(n);
    for (uint32_t t = 0; t < n; ++t) tw[t] = TrialWork{d_tim, accs[t], mean, sd, 0.0f, 0};
    psx::batch::StageTimes st;
    eng->profile(tw.data(), n, reps, st);
    const int ns = (int)st.ms.size();
    for (int k = 0; k < ns && k < *nstage; ++k) stage_ms[k] = st.ms[k];
    *nstage = ns;
    if (desc && cap) {
      strncpy(desc, st.desc.c_str(), cap - 1);
      desc[cap - 1] = 0;
    }
    eng.reset();
    (void)hipFree(d_tim);
    (void)hipStreamDestroy(s);
    return psx::ok();
  } catch (const std::exception& e) {
    if (d_tim) (void)hipFree(d_tim);
    if (s) (void)hipStreamDestroy(s);
    return psx::fail(PSX_ERR_HIP, "psx_batch_profile: %s", e.what());
  }
}

extern "C" int psx_batch_bound_audit(const psx_search_args* args, const psx_batch_opts* opts, const float* d_series,
                                     uint64_t nsamps, const float* accs, uint32_t naccs, uint64_t* result) {
  int rc = psx::search::validate_args(args);
  if (rc) return rc;
  PSX_REQUIRE(d_series && accs && naccs > 0 && result, "psx_batch_bound_audit: bad arguments");
  hipStream_t s = nullptr;
  float* d_tim = nullptr;
  try {
    EngineOptions eo;
    psx::fft::Decomp forced;
    engine_options(opts, args->size, eo, forced);
    PSX_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::unique_ptr<Engine> eng(new Engine(*args, eo, s));
    PSX_CHECK_HIP(hipMalloc(&d_tim, sizeof(float) * args->size));
    float mean, sd;
    eng->prologue(d_series, nsamps, d_tim, &mean, &sd, nullptr);
    const uint32_t n = std::min(naccs, eng->batch());
    std::vector<TrialWork> tw(n);
    for (uint32_t t = 0; t < n; ++t) tw[t] = TrialWork{d_tim, accs[t], mean, sd, 0.0f, 0};
    psx::batch::BoundAudit a;
    eng->audit_bounds(tw.data(), n, a);
    result[0] = a.entries;
    result[1] = a.mismatches;
    result[2] = a.fundamentals;
    result[3] = a.appended;
    result[4] = a.crossing;
    result[5] = a.window;
    result[6] = a.loose;
    eng.reset();
    (void)hipFree(d_tim);
    (void)hipStreamDestroy(s);
    return psx::ok();
  } catch (const std::exception& e) {
    if (d_tim) (void)hipFree(d_tim);
    if (s) (void)hipStreamDestroy(s);
    return psx::fail(PSX_ERR_HIP, "psx_batch_bound_audit: %s", e.what());
  }
}

extern "C" int psx_batch_search(const psx_search_args* args, const psx_batch_opts* opts, const float* series,
                                int series_on_device, uint64_t nsamps, const float* dm_list, uint32_t ndm,
                                const float* accs, uint32_t naccs, psx_candidate* out, uint32_t cap, uint32_t* count,
                                double* seconds) {
  int rc = psx::search::validate_args(args);
  if (rc) return rc;
  PSX_REQUIRE(series && dm_list && ndm > 0 && nsamps > 0, "psx_batch_search: bad series");
  PSX_REQUIRE(!accs || naccs > 0, "psx_batch_search: empty acceleration list");
  hipStream_t s = nullptr, ps = nullptr;
  float* d_tims = nullptr;
  float* d_stage = nullptr;
  psx::batch::PrologueResult* h_res = nullptr;
  std::vector<hipEvent_t> done_ev;
  auto release = [&]() {
    if (ps) (void)hipStreamSynchronize(ps);
    for (auto& e : done_ev)
      if (e) (void)hipEventDestroy(e);
    if (h_res) (void)hipHostFree(h_res);
    if (d_tims) (void)hipFree(d_tims);
    if (d_stage) (void)hipFree(d_stage);
    if (ps) (void)hipStreamDestroy(ps);
    if (s) (void)hipStreamDestroy(s);
  };
  try {
    EngineOptions eo;
    psx::fft::Decomp forced;
    engine_options(opts, args->size, eo, forced);
    const bool keep_engine = opts && opts->keep_engine;
    const EngineKey key = engine_key(*args, eo);
    const uint64_t N = args->size;
    std::unique_ptr<Engine> eng;
    if (!keep_engine || !take_kept(key, eng, s)) {
      PSX_CHECK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      eng.reset(new Engine(*args, eo, s));
    }
    PSX_CHECK_HIP(hipStreamCreateWithFlags(&ps, hipStreamNonBlocking));
    // declared after the engine, so destroyed first on every exit path: no
    // prologue work may still be queued on ps when the engine frees the
    // buffers it uses (a DM without trials is never waited for)
    struct DrainGuard {
      hipStream_t st;
      ~DrainGuard() { (void)hipStreamSynchronize(st); }
    } drain{ps};
    PSX_CHECK_HIP(hipMalloc(&d_tims, sizeof(float) * N * ndm));
    const uint64_t nreal = nsamps;
    auto t0 = clk::now();
    std::vector<const float*> tims(ndm);
    std::vector<float> means(ndm), stds(ndm), dms(dm_list, dm_list + ndm);
    std::vector<int> idx(ndm);
    std::vector<std::vector<float>> acc_lists(ndm);
    AccelerationPlan plan(args->acc_start, args->acc_end, args->acc_tol, args->acc_pulse_width * 1e-6, N,
                          args->tsamp, args->cfreq, args->ch_bw);
    if (!series_on_device) PSX_CHECK_HIP(hipMalloc(&d_stage, sizeof(float) * nreal));
    std::unique_ptr<Zapper> zap;
    if (args->nbirdies) zap.reset(new Zapper(args->birdie_freq, args->birdie_width, args->nbirdies, ps));
    PSX_CHECK_HIP(hipHostMalloc((void**)&h_res, sizeof(psx::batch::PrologueResult) * ndm, hipHostMallocDefault));
    done_ev.assign(ndm, nullptr);
    for (auto& e : done_ev) PSX_CHECK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    // every prologue is enqueued at once on ps; the trials of DM d wait for
    // its event only
    auto input_of = [&](uint32_t d, hipStream_t st) {
      const float* src = series + (uint64_t)d * nreal;
      if (series_on_device) return src;
      hip_check(hipMemcpyAsync(d_stage, src, sizeof(float) * nreal, hipMemcpyHostToDevice, st), "series upload");
      return (const float*)d_stage;
    };
    uint64_t ntrials = 0;
    for (uint32_t d = 0; d < ndm; ++d) {
      if (accs)
        acc_lists[d].assign(accs, accs + naccs);
      else
        plan.generate_accel_list(dms[d], args->cdm, acc_lists[d]);
      ntrials += acc_lists[d].size();
    }
    // the launch graphs of the full and the last batch exist before any
    // prologue is queued: building one later waits for the queued prologues
    eng->prepare((uint32_t)std::min<uint64_t>(eng->batch(), ntrials));
    if (ntrials > eng->batch()) eng->prepare((uint32_t)(ntrials % eng->batch()));
    {
      psx::TraceRange dm_range("DM-Loop");  // the per-DM prologues, all queued here
      for (uint32_t d = 0; d < ndm; ++d) {
        float* tim = d_tims + (uint64_t)d * N;
        eng->prologue_async(input_of(d, ps), nreal, tim, &h_res[d], ps, zap.get());
        PSX_CHECK_HIP(hipEventRecord(done_ev[d], ps));
        tims[d] = tim;
        idx[d] = (int)d;
      }
    }
    auto ready = [&](size_t d) {
      hip_check(hipEventSynchronize(done_ev[d]), "prologue event");
      if (!h_res[d].base.done) {
        // the baseline loop needs more than the device steps: redo this DM
        // with the host-driven prologue once ps is idle (its buffers are
        // shared), from its original input
        hip_check(hipStreamSynchronize(ps), "prologue stream");
        eng->prologue(input_of((uint32_t)d, s), nreal, d_tims + d * N, &means[d], &stds[d], zap.get());
        hip_check(hipStreamSynchronize(s), "prologue");
      } else {
        eng->stats_of(h_res[d], &means[d], &stds[d]);
      }
      hip_check(hipStreamWaitEvent(s, done_ev[d], 0), "stream wait");
    };
    auto t1 = clk::now();
    std::vector<std::vector<Candidate>> per_dm;
    {
      psx::TraceRange acc_range("Acceleration-Loop");  // every (DM, acceleration) trial, batched
      psx::search::run_trials(*eng, *args, tims, means, stds, dms, idx, acc_lists, per_dm, ready);
    }
    auto t2 = clk::now();
    std::vector<Candidate> flat;
    for (auto& v : per_dm) flat.insert(flat.end(), v.begin(), v.end());
    if (seconds) {
      seconds[0] = secs(t0, t1);
      seconds[1] = secs(t1, t2);
      seconds[2] = secs(t0, t2);
    }
    hip_check(hipStreamSynchronize(ps), "prologue stream");
    if (keep_engine) {
      keep(std::move(eng), s, key);
      s = nullptr;
    } else {
      eng.reset();
    }
    release();
    return psx::search::flatten(flat, out, cap, count) == PSX_OK ? psx::ok() : PSX_ERR_OVERFLOW;
  } catch (const std::exception& e) {
    release();
    return psx::fail(PSX_ERR_HIP, "psx_batch_search: %s", e.what());
  }
}
