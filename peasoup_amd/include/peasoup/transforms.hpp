// transforms.hpp - the per-DM-trial search transforms with the reference's
// class names, constructors and methods, backed by libpsx kernels:
//   TimeDomainResampler   include/transforms/resampler.hpp:7-34
//   CuFFTerR2C / C2R      include/transforms/ffter.hpp:7-110
//   SpectrumFormer        include/transforms/spectrumformer.hpp:6-24
//   Dereddener            include/transforms/dereddener.hpp:10-68
//   Zapper                include/transforms/birdiezapper.hpp:11-73
//   HarmonicFolder        include/transforms/harmonicfolder.hpp:8-45
//   PeakFinder            include/transforms/peakfinder.hpp:11-95
//   stats::*              include/utils/stats.hpp:6-43
//   AccelerationPlan      include/utils/utils.hpp:143-233
// Every class accepts an optional stream (the reference uses the legacy
// default stream, which is the default here as well).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "peasoup/candidates.hpp"
#include "peasoup/containers.hpp"

#ifndef MAX_BLOCKS
#define MAX_BLOCKS 65535
#endif
#ifndef MAX_THREADS
#define MAX_THREADS 512
#endif

// --------------------------------------------------------- resampler
class TimeDomainResampler {
 private:
  psx_stream stream;

 public:
  // launch knobs are accepted for interface compatibility; launch shapes are
  // chosen by the library.
  explicit TimeDomainResampler(unsigned int /*max_threads*/ = MAX_THREADS, unsigned int /*max_blocks*/ = MAX_BLOCKS,
                               psx_stream s = nullptr)
      : stream(s) {}
  void resample(DeviceTimeSeries<float>& input, DeviceTimeSeries<float>& output, unsigned int size, float acc) {
    peasoup::check(psx_resample_v1(input.get_data(), output.get_data(), size, acc, input.get_tsamp(), stream),
                   "Error from device_resample");
  }
  void resampleII(DeviceTimeSeries<float>& input, DeviceTimeSeries<float>& output, unsigned int size, float acc) {
    peasoup::check(psx_resample(input.get_data(), output.get_data(), size, acc, input.get_tsamp(), stream),
                   "Error from device_resampleII");
  }
};

// ------------------------------------------------------------- FFTs
class CuFFTer {
 protected:
  psx_fft_plan* fft_plan = nullptr;
  unsigned int size = 0;
  psx_stream stream = nullptr;

 public:
  CuFFTer(const CuFFTer&) = delete;
  CuFFTer& operator=(const CuFFTer&) = delete;
  CuFFTer() {}
  double get_resolution(float tsamp) { return (double)1.0 / (size * tsamp); }
  virtual unsigned int get_output_size() { return size / 2 + 1; }
  void set_stream(psx_stream s) { stream = s; }
  virtual ~CuFFTer() {
    if (fft_plan) psx_fft_plan_destroy(fft_plan);
  }
};

class CuFFTerR2C : public CuFFTer {
 public:
  explicit CuFFTerR2C(unsigned int size, long long batch = 1) {
    this->size = size;
    peasoup::check(psx_fft_plan_create(size, PSX_FFT_R2C, (uint32_t)batch, &fft_plan), "CuFFTerR2C");
  }
  void execute(float* tim, psx_c32* fseries) { peasoup::check(psx_fft_exec(fft_plan, tim, fseries, stream), "R2C"); }
};

class CuFFTerC2R : public CuFFTer {
 public:
  explicit CuFFTerC2R(unsigned int size, long long batch = 1) {
    this->size = size;
    peasoup::check(psx_fft_plan_create(size, PSX_FFT_C2R, (uint32_t)batch, &fft_plan), "CuFFTerC2R");
  }
  void execute(psx_c32* input, float* output) { peasoup::check(psx_fft_exec(fft_plan, input, output, stream), "C2R"); }
};

// --------------------------------------------------- spectrum former
class SpectrumFormer {
 private:
  psx_stream stream;

 public:
  explicit SpectrumFormer(psx_stream s = nullptr) : stream(s) {}
  void form_interpolated(DeviceFourierSeries<psx_c32>& input, DevicePowerSpectrum<float>& output) {
    peasoup::check(psx_form_interbin(input.get_data(), output.get_data(), input.get_nbins(), stream),
                   "Error from device_form_power_series");
  }
  void form(DeviceFourierSeries<psx_c32>& input, DevicePowerSpectrum<float>& output) {
    peasoup::check(psx_form_amplitude(input.get_data(), output.get_data(), input.get_nbins(), stream),
                   "Error from device_form_power_series");
  }
};

// --------------------------------------------------------- dereddener
class Dereddener {
 private:
  unsigned int size;
  float* median;
  float* work;
  psx_stream stream;

 public:
  explicit Dereddener(unsigned int size, psx_stream s = nullptr) : size(size), stream(s) {
    Utils::device_malloc(&median, size);
    Utils::device_malloc(&work, psx_deredden_work_floats(size));
  }
  Dereddener(const Dereddener&) = delete;
  Dereddener& operator=(const Dereddener&) = delete;
  ~Dereddener() {
    psx_free(median);
    psx_free(work);
  }
  // dereddener.hpp:41-62
  void calculate_median(DevicePowerSpectrum<float>& powers, float boundary_5_freq = 0.05,
                        float boundary_25_freq = 0.5) {
    if (powers.get_nbins() != size) peasoup::throw_error("Bad data length given to running_median()");
    peasoup::check(psx_deredden_median(powers.get_data(), size, powers.get_bin_width(), boundary_5_freq,
                                       boundary_25_freq, median, work, stream),
                   "calculate_median");
  }
  // dereddener.hpp:64-66
  void deredden(DeviceFourierSeries<psx_c32>& spectrum) {
    peasoup::check(psx_divide_c_by_f(spectrum.get_data(), median, spectrum.get_nbins(), stream), "deredden");
  }
  float* get_median() { return median; }
};

// ------------------------------------------------------ birdie zapper
// A zap list is lines of "<frequency Hz> <half width Hz>"; blank lines are
// skipped (birdiezapper.hpp:35-57).  A line with one field is an error here
// (the reference reads past its split result there).
class Zapper {
 private:
  std::vector<float> birdies;
  std::vector<float> widths;
  float* d_birdies = nullptr;
  float* d_widths = nullptr;
  psx_stream stream;

  void upload() {
    psx_free(d_birdies);
    psx_free(d_widths);
    d_birdies = d_widths = nullptr;
    if (birdies.empty()) return;
    Utils::device_malloc(&d_birdies, birdies.size());
    Utils::device_malloc(&d_widths, widths.size());
    Utils::h2dcpy(d_birdies, birdies.data(), birdies.size(), stream);
    Utils::h2dcpy(d_widths, widths.data(), widths.size(), stream);
  }

 public:
  explicit Zapper(const std::string& zaplist, psx_stream s = nullptr) : stream(s) { append_from_file(zaplist); }
  Zapper(const float* freqs, const float* half_widths, std::size_t n, psx_stream s = nullptr)
      : birdies(freqs, freqs + n), widths(half_widths, half_widths + n), stream(s) {
    upload();
  }
  Zapper(const Zapper&) = delete;
  Zapper& operator=(const Zapper&) = delete;
  ~Zapper() {
    psx_free(d_birdies);
    psx_free(d_widths);
  }

  void append_from_file(const std::string& zaplist) {
    std::ifstream infile(zaplist.c_str());
    if (!infile.good()) peasoup::throw_error("Zapper: cannot open " + zaplist);
    std::string line;
    while (std::getline(infile, line)) {
      std::istringstream fields(line);
      std::string f, w;
      if (!(fields >> f)) continue;
      if (!(fields >> w)) peasoup::throw_error("Zapper: line without a width in " + zaplist + ": " + line);
      birdies.push_back((float)std::atof(f.c_str()));
      widths.push_back((float)std::atof(w.c_str()));
    }
    upload();
  }

  void zap(DeviceFourierSeries<psx_c32>& fseries) {
    zap(fseries.get_data(), fseries.get_bin_width(), fseries.get_nbins());
  }
  void zap(psx_c32* fseries, float bin_width, unsigned int nbins) {
    peasoup::check(psx_zap_birdies(fseries, d_birdies, d_widths, (uint32_t)birdies.size(), bin_width, nbins, stream),
                   "Zapper::zap");
  }
  void zap_on(psx_c32* fseries, float bin_width, uint64_t nbins, psx_stream s) const {
    peasoup::check(psx_zap_birdies(fseries, d_birdies, d_widths, (uint32_t)birdies.size(), bin_width, nbins, s),
                   "Zapper::zap");
  }
  std::size_t size() const { return birdies.size(); }
  const std::vector<float>& get_birdies() const { return birdies; }
  const std::vector<float>& get_widths() const { return widths; }
};

// ------------------------------------------------------ harmonic folder
class HarmonicFolder {
 private:
  HarmonicSums<float>& sums;
  psx_stream stream;

 public:
  explicit HarmonicFolder(HarmonicSums<float>& sums, unsigned int /*max_blocks*/ = MAX_BLOCKS,
                          unsigned int /*max_threads*/ = MAX_THREADS, psx_stream s = nullptr)
      : sums(sums), stream(s) {}
  // harmonicfolder.hpp:26-38; fold pointers travel as kernel arguments
  void fold(DevicePowerSpectrum<float>& fold0) {
    float* ptrs[5];
    const uint32_t n = (uint32_t)sums.size();
    if (n < 1 || n > 5) peasoup::throw_error("HarmonicFolder supports 1..5 folds");
    for (uint32_t ii = 0; ii < n; ii++) ptrs[ii] = sums[ii]->get_data();
    peasoup::check(psx_harmonic_sum(fold0.get_data(), ptrs, n, fold0.get_nbins(), stream), "Harmonic summing");
  }
};

// --------------------------------------------------------- peak finder
namespace peasoup {
// PeakFinder::identify_unique_peaks (peakfinder.hpp:27-56): a cluster runs
// while the next index is within min_gap of the index that last set a new
// maximum.
inline int identify_unique_peaks(const int* idxs, const float* snrs, int count, int min_gap, int* peakidxs,
                                 float* peaksnrs) {
  int ii = 0, npeaks = 0;
  while (ii < count) {
    float cpeak = snrs[ii];
    int cpeakidx = idxs[ii];
    int lastidx = idxs[ii];
    ii++;
    while (ii < count && (idxs[ii] - lastidx) < min_gap) {
      if (snrs[ii] > cpeak) {
        cpeak = snrs[ii];
        cpeakidx = idxs[ii];
        lastidx = idxs[ii];
      }
      ii++;
    }
    peakidxs[npeaks] = cpeakidx;
    peaksnrs[npeaks] = cpeak;
    npeaks++;
  }
  return npeaks;
}

// PeakFinder::find_candidates search window (peakfinder.hpp:78-83).
struct PeakWindow {
  int start;
  int end;  // exclusive (min(nbins, max_bin))
  float factor;  // bin -> frequency (peakfinder.hpp:89)
};
inline PeakWindow peak_window(unsigned int nbins, double bin_width, unsigned int nh_u, float min_freq, float max_freq) {
  const int size = (int)nbins;
  const float nyquist = bin_width * size;
  const int orig_size = 2.0 * (size - 1.0);
  const int nh = (int)nh_u;
  const int max_bin = (int)((max_freq / bin_width) * std::pow(2.0, nh));
  const int start_idx = (int)(orig_size * (min_freq / nyquist) * std::pow(2.0, nh));
  PeakWindow w;
  w.start = start_idx;
  w.end = std::min(size, max_bin);
  w.factor = 1.0 / size * nyquist / std::pow(2.0, (float)nh);
  return w;
}
}  // namespace peasoup

class PeakFinder {
 private:
  float threshold;
  float min_freq;
  float max_freq;
  int min_gap;
  unsigned int max_cands;
  std::vector<int> idxs, peakidxs;
  std::vector<float> snrs, peaksnrs, peakfreqs;
  int32_t* d_idxs = nullptr;
  float* d_snrs = nullptr;
  uint32_t* d_count = nullptr;
  void* d_work = nullptr;
  psx_stream stream;

 public:
  PeakFinder(float threshold, float min_freq, float max_freq, unsigned int size, int min_gap = 30,
             psx_stream s = nullptr)
      : threshold(threshold), min_freq(min_freq), max_freq(max_freq), min_gap(min_gap), max_cands(100000),
        stream(s) {
    idxs.resize(max_cands);
    snrs.resize(max_cands);
    peakidxs.resize(max_cands);
    peaksnrs.resize(max_cands);
    peakfreqs.resize(max_cands);
    Utils::device_malloc(&d_idxs, max_cands);
    Utils::device_malloc(&d_snrs, max_cands);
    Utils::device_malloc(&d_count, 1);
    peasoup::check(psx_malloc(&d_work, psx_find_peaks_work_bytes(size / 2 + 1)), "PeakFinder");
  }
  PeakFinder(const PeakFinder&) = delete;
  PeakFinder& operator=(const PeakFinder&) = delete;
  ~PeakFinder() {
    psx_free(d_idxs);
    psx_free(d_snrs);
    psx_free(d_count);
    psx_free(d_work);
  }

  void find_candidates(HarmonicSums<float>& sums, SpectrumCandidates& cands) {
    for (std::size_t ii = 0; ii < sums.size(); ii++) find_candidates(*sums[(int)ii], cands);
  }

  // peakfinder.hpp:77-94
  void find_candidates(DevicePowerSpectrum<float>& pspec, SpectrumCandidates& cands) {
    const peasoup::PeakWindow w =
        peasoup::peak_window(pspec.get_nbins(), pspec.get_bin_width(), pspec.get_nh(), min_freq, max_freq);
    int count = 0;
    if (w.end > w.start) {
      peasoup::check(psx_find_peaks(pspec.get_data(), w.start, w.end, threshold, d_idxs, d_snrs, max_cands, d_count,
                                    d_work, stream),
                     "Error from device_find_peaks");
      uint32_t c = 0;
      Utils::d2hcpy(&c, d_count, 1, stream);
      if (c > max_cands) peasoup::throw_error("PeakFinder: more than 100000 bins above threshold");
      count = (int)c;
      if (count) {
        Utils::d2hcpy(idxs.data(), d_idxs, count, stream);
        Utils::d2hcpy(snrs.data(), d_snrs, count, stream);
      }
    }
    const int npeaks =
        peasoup::identify_unique_peaks(idxs.data(), snrs.data(), count, min_gap, peakidxs.data(), peaksnrs.data());
    for (int ii = 0; ii < npeaks; ii++) peakfreqs[ii] = peakidxs[ii] * w.factor;
    cands.append(peaksnrs.data(), peakfreqs.data(), (int)pspec.get_nh(), npeaks);
  }
};

// -------------------------------------------------------------- stats
namespace stats {
// stats.hpp:25-33 (GPU_rms, GPU_mean); work holds psx_reduce_work_bytes()
inline void stats_with(float* ptr, unsigned int nsamps, float* mean_, float* rms_, float* std_, void* work,
                       unsigned int first_samp = 0, psx_stream s = nullptr) {
  float m = 0, r = 0;
  peasoup::check(psx_mean_rms(ptr, nsamps, first_samp, &m, &r, work, s), "stats");
  *rms_ = r;
  *mean_ = m;
  *std_ = std::sqrt(r * r - m * m);
}

// Reduction scratch of one call (the reference's Thrust reductions allocate
// theirs per call as well, kernels.cu:440-455).
struct ReduceWork {
  void* p = nullptr;
  ReduceWork() { peasoup::check(psx_malloc(&p, psx_reduce_work_bytes()), "stats"); }
  ~ReduceWork() { psx_free(p); }
  ReduceWork(const ReduceWork&) = delete;
  ReduceWork& operator=(const ReduceWork&) = delete;
};

template <class T>
void stats(T* ptr, unsigned int nsamps, float* mean_, float* rms_, float* std_, unsigned int first_samp = 0) {
  ReduceWork w;
  stats_with(ptr, nsamps, mean_, rms_, std_, w.p, first_samp);
}

// stats.hpp:9-19: rms and mean of samples [first_samp, nsamps)
template <class T>
float rms(T* ptr, unsigned int nsamps, unsigned int first_samp = 0) {
  float m, r, sd;
  stats(ptr, nsamps, &m, &r, &sd, first_samp);
  return r;
}

template <class T>
float mean(T* ptr, unsigned int nsamps, unsigned int first_samp = 0) {
  float m, r, sd;
  stats(ptr, nsamps, &m, &r, &sd, first_samp);
  return m;
}

inline float std(float mean, float rms) { return std::sqrt(rms * rms - mean * mean); }

// stats.hpp:35-40
inline void normalise(float* ptr, float mean, float std_, unsigned int size, unsigned int = MAX_BLOCKS,
                      unsigned int = MAX_THREADS, psx_stream s = nullptr) {
  peasoup::check(psx_normalise(ptr, mean, std_, size, s), "Error from device_normalise");
}
}  // namespace stats

// ------------------------------------------------- acceleration plan
// include/utils/utils.hpp:143-233 (float arithmetic with double literals).
class AccelerationPlan {
 private:
  float acc_lo, acc_hi, tol, pulse_width;
  std::size_t nsamps;
  float tsamp, cfreq, ch_bw;

 public:
  AccelerationPlan(float acc_lo, float acc_hi, float tol, float pulse_width, std::size_t nsamps, float tsamp,
                   float cfreq, float ch_bw)
      : acc_lo(acc_lo), acc_hi(acc_hi), tol(tol), pulse_width(pulse_width), nsamps(nsamps), tsamp(tsamp),
        cfreq(cfreq), ch_bw(std::fabs(ch_bw)) {}

  float dispersive_smear(float dm, float cdm) const {
    float ftop = (cfreq + ch_bw / 2) / 1e9;
    float fbottom = (cfreq - ch_bw / 2) / 1e9;
    float dm_diff = std::abs(cdm - dm);
    return 4.15e-3 * (1 / (fbottom * fbottom) - 1 / (ftop * ftop)) * dm_diff;
  }
  float pulse_broadening(float dm, float cdm) const {
    float tdm = dispersive_smear(dm, cdm);
    return std::sqrt((float)((tdm * tdm) + (pulse_width * pulse_width) + (tsamp * tsamp)));
  }
  float accel_step(float dm, float cdm) const {
    float broadening = pulse_broadening(dm, cdm);
    float factor = std::sqrt((double)(tol * tol) - 1.0);
    float tobs = (float)nsamps * tsamp;
    return 2.0 * broadening * 24.0 * 299792458.0 / (double)(tobs * tobs) * factor;
  }
  void generate_accel_list(float dm, float cdm, std::vector<float>& acc_list) const {
    acc_list.clear();
    if (acc_hi == acc_lo) {
      acc_list.push_back(acc_hi);
      return;
    }
    if ((acc_hi >= 0) && (acc_lo <= 0)) acc_list.push_back(0.0f);
    float step = accel_step(dm, cdm);
    float acc = acc_lo;
    while (acc <= acc_hi) {
      acc_list.push_back(acc);
      acc += step;
    }
    acc_list.push_back(acc_hi);
    std::sort(acc_list.begin(), acc_list.end());
    acc_list.erase(std::unique(acc_list.begin(), acc_list.end()), acc_list.end());
  }
};
