// distiller.hpp - candidate distillers with the semantics of
// include/transforms/distiller.hpp:16-196 (host code, unchanged behaviour):
// candidates are sorted by descending S/N with std::sort (not stable, as in
// the reference), then each surviving candidate removes the lower-S/N ones
// related to it by the subclass condition.
#pragma once
#include <algorithm>
#include <cmath>
#include <vector>

#include "peasoup/candidates.hpp"

namespace peasoup {
struct SnrDescending {
  bool operator()(const Candidate& x, const Candidate& y) const { return x.snr > y.snr; }
};
}  // namespace peasoup

class BaseDistiller {
 protected:
  std::vector<bool> unique;
  int size = 0;
  bool keep_related;
  virtual void condition(std::vector<Candidate>& cands, int idx) = 0;
  explicit BaseDistiller(bool keep_related) : keep_related(keep_related) {}

 public:
  virtual ~BaseDistiller() {}
  // distiller.hpp:26-59
  std::vector<Candidate> distill(std::vector<Candidate>& cands) {
    size = (int)cands.size();
    unique.assign(size, true);
    std::sort(cands.begin(), cands.end(), peasoup::SnrDescending());
    int start = 0;
    while (true) {
      int idx = -1;
      for (int ii = start; ii < size; ii++) {
        if (unique[ii]) {
          start = ii + 1;
          idx = ii;
          break;
        }
      }
      if (idx == -1) break;
      condition(cands, idx);
    }
    std::vector<Candidate> out;
    for (int ii = 0; ii < size; ii++)
      if (unique[ii]) out.push_back(cands[ii]);
    return out;
  }
};

// distiller.hpp:63-108: harmonic relations k*f/(j*f0) within (1 +- tol),
// j <= max_harm, k <= 2^nh when fractional harmonics are enabled.
class HarmonicDistiller : public BaseDistiller {
 private:
  float tolerance;
  float max_harm;
  bool fractional_harms;

  void condition(std::vector<Candidate>& cands, int idx) override {
    const double upper_tol = 1 + tolerance;  // float sum, as in the reference
    const double lower_tol = 1 - tolerance;
    const double fundi_freq = cands[idx].freq;
    for (int ii = idx + 1; ii < size; ii++) {
      const double freq = cands[ii].freq;
      const int nh = cands[ii].nh;
      const float max_denominator = fractional_harms ? (float)std::pow(2.0, nh) : 1.0f;
      for (int jj = 1; jj <= max_harm; jj++) {
        for (int kk = 1; kk <= max_denominator; kk++) {
          const double ratio = kk * freq / (jj * fundi_freq);
          if (ratio > lower_tol && ratio < upper_tol) {
            if (keep_related) cands[idx].append(cands[ii]);
            unique[ii] = false;
          }
        }
      }
    }
  }

 public:
  HarmonicDistiller(float tol, float max_harm, bool keep_related, bool fractional_harms = true)
      : BaseDistiller(keep_related), tolerance(tol), max_harm(max_harm), fractional_harms(fractional_harms) {}
};

// distiller.hpp:115-163: removes candidates inside the frequency drift
// window implied by the acceleration difference.
class AccelerationDistiller : public BaseDistiller {
 private:
  float tobs;
  double tobs_over_c;
  float tolerance;

  // distiller.hpp:121-123 (returns float)
  float correct_for_acceleration(double freq, double delta_acc) { return freq + delta_acc * freq * tobs_over_c; }

  void condition(std::vector<Candidate>& cands, int idx) override {
    const double fundi_freq = cands[idx].freq;
    const double fundi_acc = cands[idx].acc;
    const double edge = fundi_freq * tolerance;
    for (int ii = idx + 1; ii < size; ii++) {
      const double delta_acc = fundi_acc - cands[ii].acc;
      const double acc_freq = correct_for_acceleration(fundi_freq, delta_acc);
      bool related;
      if (acc_freq > fundi_freq)
        related = cands[ii].freq > fundi_freq - edge && cands[ii].freq < acc_freq + edge;
      else
        related = cands[ii].freq < fundi_freq + edge && cands[ii].freq > acc_freq - edge;
      if (related) {
        if (keep_related) cands[idx].append(cands[ii]);
        unique[ii] = false;
      }
    }
  }

 public:
  AccelerationDistiller(float tobs, float tolerance, bool keep_related)
      : BaseDistiller(keep_related), tobs(tobs), tobs_over_c(tobs / 299792458.0), tolerance(tolerance) {}
  float get_tobs() const { return tobs; }
};

// distiller.hpp:167-196
class DMDistiller : public BaseDistiller {
 private:
  float tolerance;

  void condition(std::vector<Candidate>& cands, int idx) override {
    const double fundi_freq = cands[idx].freq;
    const double upper_tol = 1 + tolerance;
    const double lower_tol = 1 - tolerance;
    for (int ii = idx + 1; ii < size; ii++) {
      const double ratio = cands[ii].freq / fundi_freq;
      if (ratio > lower_tol && ratio < upper_tol) {
        if (keep_related) cands[idx].append(cands[ii]);
        unique[ii] = false;
      }
    }
  }

 public:
  DMDistiller(float tolerance, bool keep_related) : BaseDistiller(keep_related), tolerance(tolerance) {}
};
