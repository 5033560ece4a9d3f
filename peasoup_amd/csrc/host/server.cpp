// server.cpp - continuous-batching request loop for segment searches.
//
// Requests (a beam: ndm dedispersed series of one FFT length plus search
// arguments) arrive at any time through psx_server_submit.  One scheduler
// thread owns all device work:
//   admit     waiting requests, in arrival order, while their resident
//             footprint (ndm x N floats of dereddened series) plus any engine
//             they need fits the HBM budget; the series are copied into HBM,
//             the per-DM prologues run, and the request's (DM, accel) trials
//             join the trial queue of the engine for its shape;
//   dispatch  keeps two batches in flight per engine, filling each batch from
//             the head of that engine's queue - trials of different requests
//             share a batch.  Engines of different shapes either launch on
//             their own streams, so their work overlaps on the device
//             (PSX_DISPATCH_STREAMS), or take part in one graph launch per
//             mixed batch whose child graphs are one batch of every engine
//             with trials waiting (PSX_DISPATCH_MIXED_GRAPH);
//   collect   routes finished trials back to their requests; a request whose
//             trials are all done gets the Worker's per-DM
//             AccelerationDistiller, releases its memory and is completed.
// Each trial's arithmetic does not depend on which batch carries it, so a
// request's candidates are the same alone and in the mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "batch_engine.hpp"
#include "common.hpp"
#include "kernels_dev.hpp"
#include "peasoup/distiller.hpp"
#include "search_common.hpp"

using psx::batch::Engine;
using psx::batch::EngineOptions;
using psx::batch::TrialWork;

namespace {

using clk = std::chrono::steady_clock;

enum { ST_WAITING = 0, ST_ACTIVE = 1, ST_DONE = 2, ST_FAILED = 3 };

struct EngineKey {
  uint64_t size;
  float tsamp, min_snr, min_freq, max_freq;
  uint32_t nharmonics;
  int32_t min_gap;
  float freq_tol, max_harm;
  bool operator<(const EngineKey& o) const {
    return std::tie(size, tsamp, min_snr, min_freq, max_freq, nharmonics, min_gap, freq_tol, max_harm) <
           std::tie(o.size, o.tsamp, o.min_snr, o.min_freq, o.max_freq, o.nharmonics, o.min_gap, o.freq_tol,
                    o.max_harm);
  }
};

struct Request;

struct TrialRef {
  Request* req;
  uint32_t dm;
  uint32_t acc;
};

struct EngineState {
  std::unique_ptr<Engine> eng;
  hipStream_t stream = nullptr;
  hipEvent_t mark = nullptr;  // the engine stream's work so far (mixed dispatch waits on it)
  uint64_t bytes = 0;
  std::deque<TrialRef> queue;        // trials waiting for a batch
  std::deque<std::vector<TrialRef>> inflight;  // trials of each submitted batch
  clk::time_point last_used;
};

struct Request {
  uint64_t id;
  psx_request spec;
  std::vector<float> dm_list;
  std::vector<float> accs_override;
  std::vector<float> birdie_freq, birdie_width;  // the request's zap list (copied at submit)
  std::unique_ptr<Zapper> zap;                   // on the device while the request is admitted
  int state = ST_WAITING;
  int error = PSX_OK;
  std::string error_text;
  uint64_t bytes = 0;
  float* d_tims = nullptr;
  std::vector<float> means, stds;
  std::vector<std::vector<float>> acc_lists;
  std::vector<std::vector<std::vector<Candidate>>> per_trial;
  uint64_t remaining = 0, ntrials = 0;
  std::vector<psx_candidate> result;
  double t_submit = 0, t_admit = 0, t_done = 0;
  EngineState* es = nullptr;
};

double now_s() {
  static const clk::time_point t0 = clk::now();
  return std::chrono::duration<double>(clk::now() - t0).count();
}

}  // namespace

struct psx_server {
  psx_server_config cfg;
  std::map<uint64_t, psx_schedule> schedules;  // tuned schedule per FFT length
  std::map<uint64_t, psx::fft::Decomp> decomps;
  std::mutex mu;
  std::condition_variable cv;
  std::thread sched;
  bool stop = false;
  uint64_t next_id = 1;
  std::deque<Request*> waiting;
  std::vector<Request*> active;
  std::map<uint64_t, std::unique_ptr<Request>> all;
  std::map<EngineKey, std::unique_ptr<EngineState>> engines;
  // mixed dispatch: one stream, and the instantiated parent graph of every
  // combination of (engine, slot, batch size) launched so far
  hipStream_t mixed_stream = nullptr;
  std::map<std::vector<std::tuple<EngineState*, int, uint32_t>>, hipGraphExec_t> mixed_graphs;
  uint64_t used = 0;
  uint64_t completed = 0, failed = 0, trials_done = 0;
  int device = 0;

  void loop();
  bool admit_one(Request* r);
  uint32_t batch_for(uint64_t size) const {
    auto it = schedules.find(size);
    if (it != schedules.end() && it->second.batch) return it->second.batch;
    return cfg.batch ? cfg.batch : psx::batch::default_batch(size);
  }
  void finish(Request* r);
  void fail_request(Request* r, int code, const std::string& msg);
  EngineState* engine_for(const psx_search_args& a, bool create, uint64_t* need);
  void evict_idle(uint64_t need);
  bool dispatch_mixed();
  void drop_mixed_graphs(const EngineState* es);
};

namespace {

EngineKey key_of(const psx_search_args& a) {
  return EngineKey{a.size, a.tsamp, a.min_snr, a.min_freq, a.max_freq, a.nharmonics, a.min_gap, a.freq_tol,
                   a.max_harm};
}

uint64_t request_bytes(const psx_request& r) { return (uint64_t)r.ndm * r.args.size * sizeof(float); }

}  // namespace

EngineState* psx_server::engine_for(const psx_search_args& a, bool create, uint64_t* need) {
  const EngineKey k = key_of(a);
  auto it = engines.find(k);
  if (it != engines.end()) {
    if (need) *need = 0;
    return it->second.get();
  }
  const uint64_t eb = Engine::device_bytes(a.size, batch_for(a.size), 1u << 20);
  if (need) *need = eb;
  if (!create) return nullptr;
  std::unique_ptr<EngineState> es(new EngineState);
  if (hipStreamCreateWithFlags(&es->stream, hipStreamNonBlocking) != hipSuccess)
    throw std::runtime_error("stream creation failed");
  if (hipEventCreateWithFlags(&es->mark, hipEventDisableTiming) != hipSuccess) {
    (void)hipStreamDestroy(es->stream);
    throw std::runtime_error("event creation failed");
  }
  EngineOptions eo;
  eo.batch = batch_for(a.size);
  eo.use_graph = cfg.use_graph != 0;
  auto dc = decomps.find(a.size);
  if (dc != decomps.end()) eo.decomp = &dc->second;
  es->eng.reset(new Engine(a, eo, es->stream));
  es->bytes = eb;
  es->last_used = clk::now();
  used += eb;
  EngineState* raw = es.get();
  engines.emplace(k, std::move(es));
  return raw;
}

// Frees engines with no queued or in-flight work, least recently used first,
// until `need` more bytes fit the budget.
void psx_server::evict_idle(uint64_t need) {
  while (used + need > cfg.hbm_budget_bytes) {
    auto victim = engines.end();
    for (auto it = engines.begin(); it != engines.end(); ++it) {
      EngineState* e = it->second.get();
      if (!e->queue.empty() || !e->inflight.empty()) continue;
      bool busy = false;
      for (Request* r : active)
        if (r->es == e) busy = true;
      if (busy) continue;
      if (victim == engines.end() || e->last_used < victim->second->last_used) victim = it;
    }
    if (victim == engines.end()) return;
    used -= victim->second->bytes;
    drop_mixed_graphs(victim->second.get());
    victim->second->eng.reset();
    (void)hipStreamDestroy(victim->second->stream);
    (void)hipEventDestroy(victim->second->mark);
    engines.erase(victim);
  }
}

void psx_server::drop_mixed_graphs(const EngineState* es) {
  for (auto it = mixed_graphs.begin(); it != mixed_graphs.end();) {
    bool uses = false;
    for (const auto& part : it->first) uses |= std::get<0>(part) == es;
    if (uses) {
      (void)hipGraphExecDestroy(it->second);
      it = mixed_graphs.erase(it);
    } else {
      ++it;
    }
  }
}

// One mixed batch: every engine with trials waiting and a free slot stages a
// batch, and one launch of a parent graph (the staged batches' graphs as
// independent children) runs them all on the mixed stream, after each
// engine's own stream work (its requests' prologues).
bool psx_server::dispatch_mixed() {
  std::vector<EngineState*> part;
  for (auto& kv : engines) {
    EngineState* es = kv.second.get();
    if (!es->queue.empty() && es->eng->in_flight() < 2) part.push_back(es);
  }
  if (part.empty()) return false;
  if (!mixed_stream && hipStreamCreateWithFlags(&mixed_stream, hipStreamNonBlocking) != hipSuccess)
    throw std::runtime_error("stream creation failed");
  std::vector<std::tuple<EngineState*, int, uint32_t>> key;
  std::vector<TrialWork> work;
  for (EngineState* es : part) {
    Engine& eng = *es->eng;
    const uint32_t n = (uint32_t)std::min<size_t>(eng.batch(), es->queue.size());
    work.clear();
    std::vector<TrialRef> refs(es->queue.begin(), es->queue.begin() + n);
    es->queue.erase(es->queue.begin(), es->queue.begin() + n);
    for (const TrialRef& t : refs) {
      Request* r = t.req;
      work.push_back(TrialWork{r->d_tims + (uint64_t)t.dm * r->spec.args.size, r->acc_lists[t.dm][t.acc],
                               r->means[t.dm], r->stds[t.dm], r->dm_list[t.dm], (int)t.dm});
    }
    const int slot = eng.stage(work.data(), n);
    es->inflight.push_back(std::move(refs));
    es->last_used = clk::now();
    key.emplace_back(es, slot, n);
  }
  auto it = mixed_graphs.find(key);
  if (it == mixed_graphs.end()) {
    if (mixed_graphs.size() >= 256) {
      for (auto& g : mixed_graphs) (void)hipGraphExecDestroy(g.second);
      mixed_graphs.clear();
    }
    hipGraph_t pg;
    if (hipGraphCreate(&pg, 0) != hipSuccess) throw std::runtime_error("hipGraphCreate failed");
    hipGraphExec_t ge = nullptr;
    hipError_t e = hipSuccess;
    for (const auto& k : key) {
      hipGraphNode_t node;
      e = hipGraphAddChildGraphNode(&node, pg, nullptr, 0,
                                    std::get<0>(k)->eng->graph_def(std::get<1>(k), std::get<2>(k)));
      if (e != hipSuccess) break;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(&ge, pg, nullptr, nullptr, 0);
    (void)hipGraphDestroy(pg);
    if (e != hipSuccess) throw std::runtime_error(std::string("mixed batch graph: ") + hipGetErrorString(e));
    it = mixed_graphs.emplace(key, ge).first;
  }
  for (EngineState* es : part) {
    if (hipEventRecord(es->mark, es->stream) != hipSuccess ||
        hipStreamWaitEvent(mixed_stream, es->mark, 0) != hipSuccess)
      throw std::runtime_error("mixed batch: stream ordering failed");
  }
  if (hipGraphLaunch(it->second, mixed_stream) != hipSuccess) throw std::runtime_error("mixed batch: launch failed");
  for (const auto& k : key) std::get<0>(k)->eng->launched(std::get<1>(k), mixed_stream);
  return true;
}

void psx_server::fail_request(Request* r, int code, const std::string& msg) {
  r->state = ST_FAILED;
  r->error = code;
  r->error_text = msg;
  r->t_done = now_s();
  r->zap.reset();
  if (r->d_tims) {
    (void)hipFree(r->d_tims);
    r->d_tims = nullptr;
    used -= r->bytes;
  }
  ++failed;
}

bool psx_server::admit_one(Request* r) {
  uint64_t need_engine = 0;
  engine_for(r->spec.args, false, &need_engine);
  const uint64_t need = r->bytes + need_engine;
  if (used + need > cfg.hbm_budget_bytes) evict_idle(need);
  if (used + need > cfg.hbm_budget_bytes) return false;
  EngineState* es = engine_for(r->spec.args, true, nullptr);
  r->es = es;
  const psx_search_args& a = r->spec.args;
  const uint64_t N = a.size;
  if (hipMalloc(&r->d_tims, r->bytes) != hipSuccess) {
    (void)hipGetLastError();
    r->d_tims = nullptr;
    return false;
  }
  used += r->bytes;
  r->t_admit = now_s();
  r->state = ST_ACTIVE;
  const uint32_t ndm = r->spec.ndm;
  r->means.resize(ndm);
  r->stds.resize(ndm);
  r->acc_lists.resize(ndm);
  r->per_trial.resize(ndm);
  AccelerationPlan plan(a.acc_start, a.acc_end, a.acc_tol, a.acc_pulse_width * 1e-6, N, a.tsamp, a.cfreq, a.ch_bw);
  const uint64_t ncopy = std::min<uint64_t>(r->spec.nsamps, N);
  const hipMemcpyKind kind = r->spec.series_on_device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
  // every prologue and acceleration list is prepared and checked before any
  // trial of the request is queued, so a request that fails here leaves no
  // work behind
  if (!r->birdie_freq.empty())
    r->zap.reset(new Zapper(r->birdie_freq.data(), r->birdie_width.data(), r->birdie_freq.size(), es->stream));
  for (uint32_t d = 0; d < ndm; ++d) {
    float* tim = r->d_tims + (uint64_t)d * N;
    if (hipMemcpyAsync(tim, r->spec.series + (uint64_t)d * r->spec.nsamps, ncopy * sizeof(float), kind,
                       es->stream) != hipSuccess)
      throw std::runtime_error("series upload failed");
    es->eng->prologue(tim, r->spec.nsamps, tim, &r->means[d], &r->stds[d], r->zap.get());
    if (!r->accs_override.empty())
      r->acc_lists[d] = r->accs_override;
    else
      plan.generate_accel_list(r->dm_list[d], a.cdm, r->acc_lists[d]);
    for (float acc : r->acc_lists[d])
      if (!(fabs(psx::accel_fact(acc, a.tsamp)) * (double)N < 1.0))
        throw std::runtime_error("acceleration " + std::to_string(acc) + " too large for the series length");
    r->per_trial[d].resize(r->acc_lists[d].size());
  }
  for (uint32_t d = 0; d < ndm; ++d) {
    for (uint32_t j = 0; j < r->acc_lists[d].size(); ++j) es->queue.push_back(TrialRef{r, d, j});
    r->remaining += r->acc_lists[d].size();
  }
  r->ntrials = r->remaining;
  active.push_back(r);
  if (r->remaining == 0) finish(r);
  return true;
}

void psx_server::finish(Request* r) {
  const float tobs = (unsigned int)r->spec.args.size * r->spec.args.tsamp;
  std::vector<Candidate> flat;
  for (uint32_t d = 0; d < r->spec.ndm; ++d) {
    std::vector<Candidate> acc_cands;
    for (auto& v : r->per_trial[d]) acc_cands.insert(acc_cands.end(), v.begin(), v.end());
    AccelerationDistiller still(tobs, r->spec.args.freq_tol, true);
    std::vector<Candidate> dd = still.distill(acc_cands);
    flat.insert(flat.end(), dd.begin(), dd.end());
  }
  r->result.resize(psx::search::forest_size(flat));
  uint32_t cnt = 0;
  psx::search::flatten(flat, r->result.data(), (uint32_t)r->result.size(), &cnt);
  r->per_trial.clear();
  r->zap.reset();
  (void)hipFree(r->d_tims);
  r->d_tims = nullptr;
  used -= r->bytes;
  r->state = ST_DONE;
  r->t_done = now_s();
  ++completed;
  active.erase(std::remove(active.begin(), active.end(), r), active.end());
}

void psx_server::loop() {
  (void)hipSetDevice(device);
  std::vector<TrialWork> work;
  std::vector<std::vector<Candidate>> got;
  while (true) {
    bool progressed = false;
    {
      std::unique_lock<std::mutex> lk(mu);
      if (stop && waiting.empty() && active.empty()) break;
      // admission, in arrival order (a request that does not fit blocks
      // later ones: no starvation of large requests)
      while (!waiting.empty() && (cfg.max_active == 0 || active.size() < cfg.max_active)) {
        Request* r = waiting.front();
        bool ok;
        try {
          ok = admit_one(r);
        } catch (const std::exception& e) {
          fail_request(r, PSX_ERR_HIP, e.what());
          active.erase(std::remove(active.begin(), active.end(), r), active.end());
          ok = true;
        }
        if (!ok) break;
        waiting.pop_front();
        progressed = true;
      }
      try {
        if (cfg.dispatch == PSX_DISPATCH_MIXED_GRAPH)
          while (dispatch_mixed()) progressed = true;
        // dispatch + collect per engine
        for (auto& kv : engines) {
          EngineState* es = kv.second.get();
          Engine& eng = *es->eng;
          while (cfg.dispatch == PSX_DISPATCH_STREAMS && !es->queue.empty() && eng.in_flight() < 2) {
            const uint32_t n = (uint32_t)std::min<size_t>(eng.batch(), es->queue.size());
            work.clear();
            std::vector<TrialRef> refs(es->queue.begin(), es->queue.begin() + n);
            es->queue.erase(es->queue.begin(), es->queue.begin() + n);
            for (const TrialRef& t : refs) {
              Request* r = t.req;
              work.push_back(TrialWork{r->d_tims + (uint64_t)t.dm * r->spec.args.size, r->acc_lists[t.dm][t.acc],
                                       r->means[t.dm], r->stds[t.dm], r->dm_list[t.dm], (int)t.dm});
            }
            eng.submit(work.data(), n);
            es->inflight.push_back(std::move(refs));
            es->last_used = clk::now();
            progressed = true;
          }
          while (!es->inflight.empty() && eng.ready_or_done()) {
            eng.collect(got);
            std::vector<TrialRef> refs = std::move(es->inflight.front());
            es->inflight.pop_front();
            for (size_t t = 0; t < refs.size(); ++t) {
              Request* r = refs[t].req;
              r->per_trial[refs[t].dm][refs[t].acc] = std::move(got[t]);
              ++trials_done;
              if (--r->remaining == 0) finish(r);
            }
            progressed = true;
          }
        }
      } catch (const std::exception& e) {
        // a device error poisons every request in flight on the device
        for (Request* r : std::vector<Request*>(active)) fail_request(r, PSX_ERR_HIP, e.what());
        active.clear();
        for (auto& kv : engines) {
          kv.second->queue.clear();
          kv.second->inflight.clear();
        }
      }
      cv.notify_all();
      if (!progressed) cv.wait_for(lk, std::chrono::microseconds(200));
    }
  }
}

extern "C" int psx_server_create(const psx_server_config* cfg, psx_server** out) {
  PSX_REQUIRE(cfg && out, "psx_server_create: null argument");
  PSX_REQUIRE(cfg->hbm_budget_bytes > 0, "psx_server_create: zero budget");
  PSX_REQUIRE(!cfg->nschedules || cfg->schedules, "psx_server_create: null schedules");
  PSX_REQUIRE(cfg->dispatch == PSX_DISPATCH_STREAMS || (cfg->dispatch == PSX_DISPATCH_MIXED_GRAPH && cfg->use_graph),
              "psx_server_create: dispatch must be PSX_DISPATCH_STREAMS, or PSX_DISPATCH_MIXED_GRAPH with graphs");
  psx_server* s = new psx_server;
  s->cfg = *cfg;
  for (uint32_t i = 0; i < cfg->nschedules; ++i) {
    const psx_schedule& sc = cfg->schedules[i];
    if (sc.batch > 1024 || sc.fft_npass < 0 || sc.fft_npass > 3) {
      delete s;
      return psx::fail(PSX_ERR_ARG, "psx_server_create: bad schedule for size %llu", (unsigned long long)sc.size);
    }
    s->schedules[sc.size] = sc;
    if (sc.fft_npass > 0) {
      psx::fft::Decomp d;
      d.npass = sc.fft_npass;
      for (int q = 0; q < 3; ++q) {
        d.L[q] = sc.fft_len[q];
        d.C[q] = sc.fft_cols[q];
      }
      s->decomps[sc.size] = d;
    }
  }
  s->cfg.schedules = nullptr;  // copied above; the caller's array is not kept
  s->cfg.nschedules = 0;
  (void)hipGetDevice(&s->device);
  s->sched = std::thread([s]() { s->loop(); });
  *out = s;
  return psx::ok();
}

extern "C" int psx_server_submit(psx_server* s, const psx_request* req, uint64_t* id) {
  PSX_REQUIRE(s && req && id, "psx_server_submit: null argument");
  int rc = psx::search::validate_args(&req->args);
  if (rc) return rc;
  PSX_REQUIRE(req->series && req->dm_list && req->ndm > 0 && req->nsamps > 0, "psx_server_submit: bad series");
  const uint64_t need = request_bytes(*req) + Engine::device_bytes(req->args.size, s->batch_for(req->args.size), 1u << 20);
  if (need > s->cfg.hbm_budget_bytes)
    return psx::fail(PSX_ERR_BUDGET, "request needs %llu bytes, budget is %llu", (unsigned long long)need,
                     (unsigned long long)s->cfg.hbm_budget_bytes);
  std::unique_ptr<Request> r(new Request);
  r->spec = *req;
  r->dm_list.assign(req->dm_list, req->dm_list + req->ndm);
  if (req->accs) r->accs_override.assign(req->accs, req->accs + req->naccs);
  if (req->args.nbirdies) {
    r->birdie_freq.assign(req->args.birdie_freq, req->args.birdie_freq + req->args.nbirdies);
    r->birdie_width.assign(req->args.birdie_width, req->args.birdie_width + req->args.nbirdies);
  }
  r->spec.args.birdie_freq = r->spec.args.birdie_width = nullptr;  // the caller's arrays are not kept
  r->bytes = request_bytes(*req);
  r->t_submit = now_s();
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return psx::fail(PSX_ERR_STATE, "psx_server_submit: server is shutting down");
  r->id = s->next_id++;
  *id = r->id;
  s->waiting.push_back(r.get());
  s->all.emplace(r->id, std::move(r));
  s->cv.notify_all();
  return psx::ok();
}

static void fill_status(const Request* r, psx_request_status* st) {
  st->state = r->state;
  st->error = r->error;
  st->ncand = (uint32_t)r->result.size();
  st->ntrials = r->ntrials;
  st->t_submit = r->t_submit;
  st->t_admit = r->t_admit;
  st->t_done = r->t_done;
}

extern "C" int psx_server_poll(psx_server* s, uint64_t id, psx_request_status* st) {
  PSX_REQUIRE(s && st, "psx_server_poll: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  PSX_REQUIRE(it != s->all.end(), "psx_server_poll: unknown request %llu", (unsigned long long)id);
  fill_status(it->second.get(), st);
  return psx::ok();
}

extern "C" int psx_server_wait(psx_server* s, uint64_t id, double timeout_s, psx_request_status* st) {
  PSX_REQUIRE(s && st, "psx_server_wait: null argument");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  PSX_REQUIRE(it != s->all.end(), "psx_server_wait: unknown request %llu", (unsigned long long)id);
  Request* r = it->second.get();
  const auto deadline = clk::now() + std::chrono::duration<double>(timeout_s);
  while (r->state != ST_DONE && r->state != ST_FAILED) {
    if (s->cv.wait_until(lk, deadline) == std::cv_status::timeout &&
        r->state != ST_DONE && r->state != ST_FAILED) {
      fill_status(r, st);
      return psx::fail(PSX_ERR_TIMEOUT, "psx_server_wait: request %llu not done", (unsigned long long)id);
    }
  }
  fill_status(r, st);
  if (r->state == ST_FAILED) return psx::fail(r->error, "request failed: %s", r->error_text.c_str());
  return psx::ok();
}

extern "C" int psx_server_result(psx_server* s, uint64_t id, psx_candidate* out, uint32_t cap, uint32_t* count) {
  PSX_REQUIRE(s && count, "psx_server_result: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  PSX_REQUIRE(it != s->all.end(), "psx_server_result: unknown request %llu", (unsigned long long)id);
  Request* r = it->second.get();
  if (r->state == ST_FAILED) return psx::fail(r->error, "request failed: %s", r->error_text.c_str());
  if (r->state != ST_DONE) return psx::fail(PSX_ERR_STATE, "psx_server_result: request not done");
  *count = (uint32_t)r->result.size();
  if (out)
    std::memcpy(out, r->result.data(), sizeof(psx_candidate) * std::min<size_t>(cap, r->result.size()));
  if (r->result.size() > cap) return psx::fail(PSX_ERR_OVERFLOW, "result holds %zu candidates", r->result.size());
  s->all.erase(it);
  return psx::ok();
}

extern "C" int psx_server_stats(psx_server* s, psx_server_counters* st) {
  PSX_REQUIRE(s && st, "psx_server_stats: null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  st->used_bytes = s->used;
  st->budget_bytes = s->cfg.hbm_budget_bytes;
  st->waiting = (uint32_t)s->waiting.size();
  st->active = (uint32_t)s->active.size();
  st->completed = s->completed;
  st->failed = s->failed;
  st->trials_done = s->trials_done;
  st->engines = (uint32_t)s->engines.size();
  return psx::ok();
}

extern "C" int psx_server_destroy(psx_server* s) {
  if (!s) return psx::ok();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->sched.joinable()) s->sched.join();
  if (s->mixed_stream) (void)hipStreamSynchronize(s->mixed_stream);
  for (auto& g : s->mixed_graphs) (void)hipGraphExecDestroy(g.second);
  for (auto& kv : s->engines) {
    kv.second->eng.reset();
    (void)hipStreamDestroy(kv.second->stream);
    (void)hipEventDestroy(kv.second->mark);
  }
  if (s->mixed_stream) (void)hipStreamDestroy(s->mixed_stream);
  for (auto& kv : s->all)
    if (kv.second->d_tims) (void)hipFree(kv.second->d_tims);
  delete s;
  return psx::ok();
}
