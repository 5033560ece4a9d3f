// worker.cpp - Worker-equivalent harness: the reference's per-GPU search loop
// (src/pipeline_multi.cu:111-303) written against the drop-in transform
// classes, one library call per reference call.  It is the "plain" form the
// batched search is measured against, and shows that the classes compose in
// the reference's order.
#include <chrono>
#include <cstring>
#include <exception>
#include <memory>
#include <vector>

#include "peasoup/candidates.hpp"
#include "peasoup/containers.hpp"
#include "peasoup/distiller.hpp"
#include "peasoup/transforms.hpp"
#include "search_common.hpp"
#include "trace.hpp"

namespace {

using clk = std::chrono::steady_clock;

double secs(clk::time_point a, clk::time_point b) { return std::chrono::duration<double>(b - a).count(); }

struct Timings {
  double prologue = 0, accel = 0, distill = 0;
};

// pipeline_multi.cu:111-303 for the DM indices of `trials`.
void run_worker(const psx_search_args& args, DispersionTrials<float>& trials, const float* accs, uint32_t naccs,
                CandidateCollection& dm_trial_cands, Timings& tm) {
  const unsigned int size = (unsigned int)args.size;
  const bool padding = size > trials.get_nsamps();
  CuFFTerR2C r2cfft(size);
  CuFFTerC2R c2rfft(size);
  const float tobs = size * trials.get_tsamp();
  const float bin_width = 1.0 / tobs;
  DeviceFourierSeries<psx_c32> d_fseries(size / 2 + 1, bin_width);
  DedispersedTimeSeries<float> tim;
  ReusableDeviceTimeSeries<float, float> d_tim(size);
  DeviceTimeSeries<float> d_tim_r(size);
  TimeDomainResampler resampler;
  DevicePowerSpectrum<float> pspec(d_fseries);
  Dereddener rednoise(size / 2 + 1);
  SpectrumFormer former;
  PeakFinder cand_finder(args.min_snr, args.min_freq, args.max_freq, size, args.min_gap);
  HarmonicSums<float> sums(pspec, args.nharmonics);
  HarmonicFolder harm_folder(sums);
  std::vector<float> acc_list;
  HarmonicDistiller harm_finder(args.freq_tol, args.max_harm, false);
  AccelerationDistiller acc_still(tobs, args.freq_tol, true);
  AccelerationPlan acc_plan(args.acc_start, args.acc_end, args.acc_tol, args.acc_pulse_width * 1e-6, size,
                            args.tsamp, args.cfreq, args.ch_bw);
  // pipeline_multi.cu:136-141: the zap list is optional
  std::unique_ptr<Zapper> bzap;
  if (args.nbirdies) bzap.reset(new Zapper(args.birdie_freq, args.birdie_width, args.nbirdies));
  void* stats_work = nullptr;
  peasoup::check(psx_malloc(&stats_work, psx_reduce_work_bytes()), "Worker");
  float mean, std, rms;
  try {
    for (std::size_t ii = 0; ii < trials.get_count(); ++ii) {
      psx::TraceRange dm_range("DM-Loop");
      auto t0 = clk::now();
      trials.get_idx(ii, tim, size);
      d_tim.copy_from_host(tim);
      d_tim.remove_baseline(std::min(tim.get_nsamps(), d_tim.get_nsamps()));
      if (padding && tim.get_nsamps() < d_tim.get_nsamps()) d_tim.fill(trials.get_nsamps(), d_tim.get_nsamps(), 0);
      if (accs)
        acc_list.assign(accs, accs + naccs);
      else
        acc_plan.generate_accel_list(tim.get_dm(), args.cdm, acc_list);
      r2cfft.execute(d_tim.get_data(), d_fseries.get_data());
      former.form(d_fseries, pspec);
      rednoise.calculate_median(pspec);
      rednoise.deredden(d_fseries);
      if (bzap) bzap->zap(d_fseries);  // pipeline_multi.cu:222-226
      former.form_interpolated(d_fseries, pspec);
      stats::stats_with(pspec.get_data(), size / 2 + 1, &mean, &rms, &std, stats_work);
      c2rfft.execute(d_fseries.get_data(), d_tim.get_data());
      auto t1 = clk::now();
      CandidateCollection accel_trial_cands;
      for (std::size_t jj = 0; jj < acc_list.size(); jj++) {
        psx::TraceRange acc_range("Acceleration-Loop");
        resampler.resampleII(d_tim, d_tim_r, size, acc_list[jj]);
        r2cfft.execute(d_tim_r.get_data(), d_fseries.get_data());
        former.form_interpolated(d_fseries, pspec);
        stats::normalise(pspec.get_data(), mean * size, std * size, size / 2 + 1);
        harm_folder.fold(pspec);
        SpectrumCandidates trial_cands(tim.get_dm(), (int)ii, acc_list[jj]);
        cand_finder.find_candidates(pspec, trial_cands);
        cand_finder.find_candidates(sums, trial_cands);
        accel_trial_cands.append(harm_finder.distill(trial_cands.cands));
      }
      auto t2 = clk::now();
      dm_trial_cands.append(acc_still.distill(accel_trial_cands.cands));
      auto t3 = clk::now();
      tm.prologue += secs(t0, t1);
      tm.accel += secs(t1, t2);
      tm.distill += secs(t2, t3);
    }
  } catch (...) {
    psx_free(stats_work);
    throw;
  }
  psx_free(stats_work);
}

}  // namespace

namespace psx {
namespace search {

int validate_args(const psx_search_args* a) {
  PSX_REQUIRE(a, "null search args");
  PSX_REQUIRE(a->size >= 64 && a->size <= (1ull << 28), "size must be in [64, 2^28]");
  PSX_REQUIRE(a->tsamp > 0, "tsamp must be > 0");
  PSX_REQUIRE(a->nharmonics >= 1 && a->nharmonics <= 5, "nharmonics must be 1..5");
  PSX_REQUIRE(a->min_gap >= 1, "min_gap must be >= 1");
  PSX_REQUIRE(!a->nbirdies || (a->birdie_freq && a->birdie_width), "birdie arrays missing");
  return PSX_OK;
}

}  // namespace search
}  // namespace psx

extern "C" int psx_worker_search(const psx_search_args* args, const float* series, uint64_t nsamps,
                                 const float* dm_list, uint32_t ndm, const float* accs, uint32_t naccs,
                                 psx_candidate* out, uint32_t cap, uint32_t* count, double* seconds) {
  int rc = psx::search::validate_args(args);
  if (rc) return rc;
  PSX_REQUIRE(series && dm_list && ndm > 0 && nsamps > 0, "psx_worker_search: bad series");
  PSX_REQUIRE(!accs || naccs > 0, "psx_worker_search: empty acceleration list");
  try {
    DispersionTrials<float> trials(args->tsamp);
    trials.resize(nsamps, std::vector<float>(dm_list, dm_list + ndm));
    std::memcpy(trials.get_data_ptr(), series, sizeof(float) * nsamps * ndm);
    CandidateCollection dm_trial_cands;
    Timings tm;
    run_worker(*args, trials, accs, naccs, dm_trial_cands, tm);
    if (seconds) {
      seconds[0] = tm.prologue;
      seconds[1] = tm.accel;
      seconds[2] = tm.distill;
    }
    return psx::search::flatten(dm_trial_cands.cands, out, cap, count) == PSX_OK ? psx::ok() : PSX_ERR_OVERFLOW;
  } catch (const std::exception& e) {
    return psx::fail(PSX_ERR_HIP, "psx_worker_search: %s", e.what());
  }
}
