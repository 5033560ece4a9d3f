// candidates.hpp - candidate records (include/data_types/candidates.hpp:10-166).
// Field layout of CandidatePOD and the Candidate fields used by the search
// path are kept as in the reference because downstream writers depend on them.
#pragma once
#include <cstdio>
#include <vector>

struct CandidatePOD {
  float dm;
  int dm_idx;
  float acc;
  int nh;
  float snr;
  float freq;
};

struct Candidate {
 public:
  float dm;
  int dm_idx;
  float acc;
  int nh;
  float snr;
  float freq;
  float folded_snr;
  double opt_period;
  bool is_adjacent;
  bool is_physical;
  float ddm_count_ratio;
  float ddm_snr_ratio;
  std::vector<Candidate> assoc;
  std::vector<float> fold;
  int nbins;
  int nints;

  Candidate(float dm, int dm_idx, float acc, int nh, float snr, float freq)
      : dm(dm), dm_idx(dm_idx), acc(acc), nh(nh), snr(snr), freq(freq), folded_snr(0.0), opt_period(0.0),
        is_adjacent(false), is_physical(false), ddm_count_ratio(0.0), ddm_snr_ratio(0.0), nbins(0), nints(0) {}
  Candidate()
      : dm(0), dm_idx(0), acc(0), nh(0), snr(0), freq(0), folded_snr(0.0), opt_period(0.0), is_adjacent(false),
        is_physical(false), ddm_count_ratio(0.0), ddm_snr_ratio(0.0), nbins(0), nints(0) {}

  void append(Candidate& other) { assoc.push_back(other); }

  // candidates.hpp:44-51
  int count_assoc() const {
    int count = 0;
    for (const Candidate& a : assoc) count += 1 + a.count_assoc();
    return count;
  }

  // candidates.hpp:71-77
  void collect_candidates(std::vector<CandidatePOD>& lite) const {
    lite.push_back(CandidatePOD{dm, dm_idx, acc, nh, snr, freq});
    for (const Candidate& a : assoc) a.collect_candidates(lite);
  }
};

class CandidateCollection {
 public:
  std::vector<Candidate> cands;
  CandidateCollection() {}
  void append(CandidateCollection& other) { cands.insert(cands.end(), other.cands.begin(), other.cands.end()); }
  void append(std::vector<Candidate> other) { cands.insert(cands.end(), other.begin(), other.end()); }
  void reset() { cands.clear(); }
};

// candidates.hpp:152-166
class SpectrumCandidates : public CandidateCollection {
 public:
  float dm;
  int dm_idx;
  float acc;
  SpectrumCandidates(float dm, int dm_idx, float acc) : dm(dm), dm_idx(dm_idx), acc(acc) {}
  void append(const float* snrs, const float* freqs, int nh, int size) {
    cands.reserve(size + cands.size());
    for (int ii = 0; ii < size; ii++) cands.push_back(Candidate(dm, dm_idx, acc, nh, snrs[ii], freqs[ii]));
  }
};
