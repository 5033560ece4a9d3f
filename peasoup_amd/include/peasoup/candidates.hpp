// candidates.hpp - candidate records (include/data_types/candidates.hpp:10-166).
// Field layout of CandidatePOD and the Candidate fields used by the search
// path are kept as in the reference because downstream writers depend on them.
#pragma once
#include <cstdio>
#include <stdexcept>
#include <string>
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
  // candidates.hpp:57-61 (folded S/N given)
  Candidate(float dm, int dm_idx, float acc, int nh, float snr, float folded_snr, float freq)
      : dm(dm), dm_idx(dm_idx), acc(acc), nh(nh), snr(snr), freq(freq), folded_snr(folded_snr), opt_period(0.0),
        is_adjacent(false), is_physical(false), ddm_count_ratio(0.0), ddm_snr_ratio(0.0), nbins(0), nints(0) {}
  Candidate()
      : dm(0), dm_idx(0), acc(0), nh(0), snr(0), freq(0), folded_snr(0.0), opt_period(0.0), is_adjacent(false),
        is_physical(false), ddm_count_ratio(0.0), ddm_snr_ratio(0.0), nbins(0), nints(0) {}

  // candidates.hpp:69-76: keep a copy of an nbins x nints fold
  void set_fold(const float* ar, int nbins_, int nints_) {
    nbins = nbins_;
    nints = nints_;
    fold.assign(ar, ar + (size_t)nbins_ * (size_t)nints_);
  }

  // candidates.hpp:86-97: one tab-separated line (period, optimal period,
  // frequency, DM, acceleration, harmonic, S/N, folded S/N, adjacency,
  // physicality, the two DM ratios, number of direct associates), then the
  // associates' lines depth first
  void print(FILE* fo = stdout) const {
    std::fprintf(fo, "%.15f\t%.15f\t%.15f\t%.2f\t%.2f\t%d\t%.1f\t%.1f\t%d\t%d\t%.4f\t%.4f\t%zu\n", 1.0 / freq,
                 opt_period, freq, dm, acc, nh, snr, folded_snr, (int)is_adjacent, (int)is_physical,
                 ddm_count_ratio, ddm_snr_ratio, assoc.size());
    for (const Candidate& a : assoc) a.print(fo);
  }

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

  // candidates.hpp:117-120
  void print(FILE* fo = stdout) const {
    for (const Candidate& c : cands) c.print(fo);
  }

  // candidates.hpp:122-138: one text file per candidate,
  // <dir>/cand_<index>_<period>_<dm>_<acc>.peasoup holding its print()
  void generate_candidate_binaries(std::string output_directory = "./") const {
    for (size_t ii = 0; ii < cands.size(); ++ii) {
      const Candidate& c = cands[ii];
      char name[128];
      std::snprintf(name, sizeof(name), "cand_%04d_%.5f_%.1f_%.1f.peasoup", (int)ii, 1.0 / c.freq, c.dm, c.acc);
      const std::string path = output_directory + "/" + name;
      FILE* fo = std::fopen(path.c_str(), "w");
      if (!fo) throw std::runtime_error("cannot open " + path);
      c.print(fo);
      std::fclose(fo);
    }
  }

  // candidates.hpp:140-148: header line, then "#Candidate <i>" and its print()
  void write_candidate_file(std::string filepath = "./candidates.txt") const {
    FILE* fo = std::fopen(filepath.c_str(), "w");
    if (!fo) throw std::runtime_error("cannot open " + filepath);
    std::fprintf(fo, "#Period...Optimal period...Frequency...DM...Acceleration...Harmonic number...S/N...Folded S/N\n");
    for (size_t ii = 0; ii < cands.size(); ++ii) {
      std::fprintf(fo, "#Candidate %d\n", (int)ii);
      cands[ii].print(fo);
    }
    std::fclose(fo);
  }
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
