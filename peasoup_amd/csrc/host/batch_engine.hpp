// batch_engine.hpp - batched acceleration search for one FFT length.
//
// The engine owns the device buffers of one batch of up to B trials and two
// sets of candidate/descriptor buffers, so the host can cluster and distill
// the candidates of batch i (PeakFinder::identify_unique_peaks,
// HarmonicDistiller - host code, semantics unchanged) while batch i+1 runs on
// the device.  Each batch is one hipGraph launch (captured once per batch
// shape) when graphs are enabled.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "batch.hpp"
#include "fft_plan.hpp"
#include "reduce_async.hpp"
#include "peasoup/candidates.hpp"
#include "peasoup/transforms.hpp"
#include "psx.h"

namespace psx {
namespace batch {

struct TrialWork {
  const float* tim;  // dereddened, inverse-transformed series of the DM (device)
  float acc;
  float mean;  // interbinned-spectrum stats of the DM (stats::stats)
  float std;
  float dm;
  int dm_idx;
};

// Results of an asynchronous prologue, written by the device into pinned
// host memory: the baseline loop's state and the interbinned spectrum's sums.
struct PrologueResult {
  BaselineState base;
  double2 sums;  // (sum x, sum x^2) over the nb bins
};

// Trials per fused batch when the caller leaves it to the library: 64 up to
// 2^26 samples (the tuned choice on MI355X from 2^20 to 2^26; 1.9% faster
// than 32 at 2^23), then halved per doubling (about 32 GB of FFT buffers).
inline uint32_t default_batch(uint64_t n) { return n <= (1ull << 26) ? 64u : (n <= (1ull << 27) ? 32u : 16u); }

struct EngineOptions {
  uint32_t batch = 16;
  uint32_t cand_capacity = 1u << 20;
  bool use_graph = true;
  const psx::fft::Decomp* decomp = nullptr;  // FFT decomposition override (autotuner)
  int spectrum_layout = PSX_SPECTRUM_AUTO;
  int harmonic_pass = PSX_HARMONIC_BOUND;
};

// Self-check of the bound spectrum one batch leaves (Engine::audit_bounds).
struct BoundAudit {
  uint64_t entries = 0;       // bound entries checked (all trials)
  uint64_t mismatches = 0;    // entries below the pair maximum of P (invalid bounds)
  uint64_t loose = 0;         // valid entries above the tight bound (bf16 rounding up of the maximum)
  uint64_t fundamentals = 0;  // bins above threshold in the fundamental's window (host scan of P)
  uint64_t appended = 0;      // candidates the device appended (the fundamental's, before the harmonic pass)
  uint64_t crossing = 0;      // bins of the harmonic window whose bound crosses at some fold (summed exactly)
  uint64_t window = 0;        // bins of the harmonic window (all trials)
};

struct StageTimes {
  // per launch: FFT passes (first one with resampling), spectrum pass (with
  // the bound pooling and the fundamental's threshold in the bound form),
  // harmonic pass
  std::vector<double> ms;
  int npass = 0;
  std::string desc;  // FFT decomposition
};

class Engine {
 public:
  Engine(const psx_search_args& args, const EngineOptions& opt, hipStream_t stream);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  uint32_t batch() const { return B; }
  uint64_t size() const { return N; }
  hipStream_t stream() const { return s; }
  // Enqueue up to B trials.  At most two batches are in flight; submit()
  // collects the oldest first when both slots are busy (its results are
  // queued for the next collect()).
  void submit(const TrialWork* trials, uint32_t n);
  // Captures and instantiates both slots' launch graphs for batches of n
  // trials now, so that the first submits do not instantiate while other
  // streams' work is queued (instantiation waits for the device).
  void prepare(uint32_t n);
  // Mixed-length dispatch (one graph launch carries a batch of several
  // engines): stage() fills the next slot with up to B trials and makes sure
  // its launch graph for n trials exists, without launching it, and returns
  // the slot; the caller launches graph_def(slot, n), as a child of its own
  // graph, on a stream of its choice and then calls launched(slot, stream),
  // which records the slot's completion there.  Needs graphs enabled.
  int stage(const TrialWork* trials, uint32_t n);
  hipGraph_t graph_def(int slot, uint32_t n) const;
  void launched(int slot, hipStream_t ls);
  // Waits for the oldest in-flight batch and returns its per-trial
  // harmonic-distilled candidate lists (trial order of the submit).
  bool collect(std::vector<std::vector<Candidate>>& per_trial);
  uint32_t in_flight() const { return (uint32_t)(pending.size() + ready.size()); }
  // True when collect() would return without waiting on the device.
  bool ready_or_done();

  // Worker prologue (pipeline_multi.cu:170-238) for one DM with the plain
  // kernels: d_series (n_real samples) -> d_tim (N samples), mean/std.
  // zap (optional) removes birdies after dereddening (pipeline_multi.cu:222-226).
  void prologue(const float* d_series, uint64_t n_real, float* d_tim, float* mean, float* std,
                const Zapper* zap = nullptr);
  // The same prologue enqueued on stream ps with no host round trip: the
  // baseline loop runs kBaselineSteps device iterations and the statistics
  // arrive in *res (pinned host memory) when ps reaches them.  Consecutive
  // calls must share ps (they reuse the prologue buffers in order).  If
  // res->base.done is 0 once ps has passed the call, the baseline loop had
  // not stopped: redo the DM with prologue().  stats_of turns *res into the
  // (mean, std) prologue() returns.
  static constexpr int kBaselineSteps = 4;
  void prologue_async(const float* d_series, uint64_t n_real, float* d_tim, PrologueResult* res, hipStream_t ps,
                      const Zapper* zap = nullptr);
  void stats_of(const PrologueResult& res, float* mean, float* std) const;

  // Average duration of each launch of one batch of n trials, timed with HIP
  // events on the engine stream over `reps` repetitions (after one warm-up).
  void profile(const TrialWork* trials, uint32_t n, int reps, StageTimes& out);
  // Runs the FFT and spectrum stages of one batch of n trials and checks the
  // bound spectrum against P on the host (bound form only).
  void audit_bounds(const TrialWork* trials, uint32_t n, BoundAudit& out);

  // Device bytes held by an engine of this shape (for admission control).
  static uint64_t device_bytes(uint64_t N, uint32_t B, uint32_t cap);

 private:
  struct Slot {
    psx::fft::TrialDesc* d_desc = nullptr;
    psx::fft::TrialDesc* h_desc = nullptr;
    Cand16* d_cands = nullptr;
    Cand16* h_cands = nullptr;
    uint32_t* d_count = nullptr;
    uint32_t* h_count = nullptr;
    hipEvent_t done = nullptr;
    std::vector<TrialWork> meta;
    std::map<uint32_t, hipGraphExec_t> graphs;
    std::map<uint32_t, hipGraph_t> defs;  // the captured graphs, for mixed dispatch
  };
  void fill_slot(Slot& sl, const TrialWork* trials, uint32_t n);
  void enqueue(Slot& sl, uint32_t n);
  // enqueue's body; run = false only builds the slot's graph for n
  void launch(Slot& sl, uint32_t n, bool run);
  void spectrum(Slot& sl, uint32_t n);
  void harmonics(Slot& sl, uint32_t n);
  // the fused FFT leaves the transposed row layout (rows_R rows of rows_L)
  bool rows_out = false;
  uint32_t rows_R = 0, rows_L = 0;
  bool collect_oldest(std::vector<std::vector<Candidate>>& per_trial);
  void process(Slot& sl, std::vector<std::vector<Candidate>>& per_trial);

  psx_search_args args;
  uint64_t N, M, nb;
  int64_t pstride;
  uint32_t B, cap, F;
  bool use_graph;
  hipStream_t s;
  psx_fft_plan* plan = nullptr;
  psx::fft::RootTable rt{};
  float2* Z = nullptr;
  float* P = nullptr;
  // bound-filtered harmonic pass: Q rows of qstride bfloat16 bounds;
  // fused: the row-layout spectrum pass writes them (and no P, which the
  // harmonic pass's exact sums recompute from Z)
  bool bound = true;
  bool fused = false;
  void* rows_src = nullptr;  // device copy of the rows source (fused form)
  int64_t qstride = 0;
  uint16_t* Q = nullptr;
  Windows win{};
  float factor[6];
  Slot slots[2];
  std::vector<int> pending;  // slot indices in submission order
  int next_slot = 0;
  std::vector<std::vector<std::vector<Candidate>>> ready;  // collected early by submit()

  // prologue resources
  psx_fft_plan* p_r2c = nullptr;
  psx_fft_plan* p_c2r = nullptr;
  psx_c32* X = nullptr;
  float* A = nullptr;
  float* med = nullptr;
  float* med_work = nullptr;
  void* red_work = nullptr;
  BaselineState* d_base = nullptr;
  double2* d_sums = nullptr;
};

}  // namespace batch
}  // namespace psx
