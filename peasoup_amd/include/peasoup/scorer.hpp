// scorer.hpp - candidate scoring of the search program
// (include/transforms/scorer.hpp:8-87, used at src/pipeline_multi.cu:550-552).
//
// Three scores per final candidate, computed from its directly associated
// candidates (the ones the DM and harmonic distillers folded into it):
//   is_physical      the period exceeds the intra-channel DM smearing at the
//                    candidate's DM (8300 foff / cfreq^3 per unit DM)
//   is_adjacent      an associated candidate sits one DM trial away, or every
//                    associated candidate shares the candidate's DM trial
//   ddm_*_ratio      share of associated candidates (by count and by summed
//                    S/N) within one "band-smearing DM width" of the
//                    candidate's DM, 1 / (freq * 4150 (fbottom^-2 - ftop^-2))
// The constants are stored as float after a double evaluation and every test
// compares float products, as the reference does, so the flags and ratios
// are the reference's bit for bit.
#pragma once
#include <cmath>
#include <cstdlib>
#include <vector>

#include "peasoup/candidates.hpp"

class CandidateScorer {
 public:
  // Units are the filterbank header's (cfreq, foff and bw in MHz).
  // tsamp is accepted for the reference's signature; no score uses it.
  CandidateScorer(float tsamp, float cfreq, float foff, float bw) {
    (void)tsamp;
    const double half = (double)bw / 2.0;
    const float f_hi = (float)((double)cfreq + half);
    const float f_lo = (float)((double)cfreq - half);
    chan_smear_per_dm_ = (float)(8300.0 * (double)foff / std::pow((double)cfreq, 3.0));
    const double inv_lo2 = 1.0 / std::pow((double)f_lo, 2.0);
    const double inv_hi2 = 1.0 / std::pow((double)f_hi, 2.0);
    band_smear_per_dm_ = (float)(4150.0 * (inv_lo2 - inv_hi2));
  }

  void score(Candidate& cand) const {
    cand.is_physical = physical(cand);
    cand.is_adjacent = adjacent(cand);
    dm_ratios(cand);
  }

  void score_all(std::vector<Candidate>& cands) const {
    for (Candidate& c : cands) score(c);
  }

 private:
  float chan_smear_per_dm_;  // seconds of channel smearing per unit DM
  float band_smear_per_dm_;  // seconds of band smearing per unit DM

  bool physical(const Candidate& c) const {
    const float smear = c.dm * chan_smear_per_dm_;
    return 1.0 / (double)c.freq > (double)smear;
  }

  // The first associated candidate one DM trial away decides; otherwise the
  // candidate is "adjacent" only if none of its associates has another DM.
  bool adjacent(const Candidate& c) const {
    bool same_dm_only = true;
    for (const Candidate& a : c.assoc) {
      if (a.dm_idx == c.dm_idx + 1 || a.dm_idx == c.dm_idx - 1) return true;
      if (a.dm_idx != c.dm_idx) same_dm_only = false;
    }
    return same_dm_only;
  }

  void dm_ratios(Candidate& c) const {
    const float width = (float)(1.0 / (double)(c.freq * band_smear_per_dm_));
    int n_in = 1, n_all = 1;
    float snr_in = c.snr, snr_all = c.snr;
    for (const Candidate& a : c.assoc) {
      ++n_all;
      snr_all += a.snr;
      if (std::fabs(c.dm - a.dm) <= width) {
        ++n_in;
        snr_in += a.snr;
      }
    }
    c.ddm_count_ratio = (float)n_in / (float)n_all;
    c.ddm_snr_ratio = snr_in / snr_all;
  }
};
