// ref_driver.cpp - C entry points over the reference's own distillers and
// candidate scorer.
//
// Compiled by oracle/ref.mk against the reference sources (include/transforms/
// distiller.hpp, include/transforms/scorer.hpp and include/data_types/
// candidates.hpp, which need only the C++ standard library) into
// oracle/_ref/libref.so.  Test infrastructure only: it checks the oracle's
// restatement of the distillers, the scorer and main()'s post-processing.
#include <algorithm>

#include "transforms/distiller.hpp"
#include "transforms/scorer.hpp"

namespace {
std::vector<Candidate> rows_to_cands(const float* rows, int n) {
  std::vector<Candidate> v;
  v.reserve(n);
  for (int i = 0; i < n; ++i) {
    const float* r = rows + 6 * i;
    v.push_back(Candidate(r[0], (int)r[1], r[2], (int)r[3], r[4], r[5]));
  }
  return v;
}
int cands_to_rows(std::vector<Candidate>& v, float* rows, int* nassoc) {
  for (size_t i = 0; i < v.size(); ++i) {
    float* r = rows + 6 * i;
    r[0] = v[i].dm;
    r[1] = (float)v[i].dm_idx;
    r[2] = v[i].acc;
    r[3] = (float)v[i].nh;
    r[4] = v[i].snr;
    r[5] = v[i].freq;
    nassoc[i] = v[i].count_assoc();
  }
  return (int)v.size();
}
// Forest rows: 6 floats per record (dm, dm_idx, acc, nh, snr, freq) and the
// record's subtree size in nassoc, pre-order.
int tree_from_rows(const float* rows, const int* nassoc, int n, int i, Candidate& c) {
  const float* r = rows + 6 * i;
  c = Candidate(r[0], (int)r[1], r[2], (int)r[3], r[4], r[5]);
  const int end = i + 1 + nassoc[i];
  if (nassoc[i] < 0 || end > n) return -1;
  int j = i + 1;
  while (j < end) {
    Candidate child;
    j = tree_from_rows(rows, nassoc, n, j, child);
    if (j < 0) return -1;
    c.append(child);
  }
  return j;
}

int forest_from_rows(const float* rows, const int* nassoc, int n, std::vector<Candidate>& v) {
  int i = 0;
  while (i < n) {
    Candidate c;
    i = tree_from_rows(rows, nassoc, n, i, c);
    if (i < 0) return -1;
    v.push_back(c);
  }
  return 0;
}

void rows_from_tree(Candidate& c, float* rows, int* nassoc, int cap, int& pos) {
  if (pos < cap) {
    float* r = rows + 6 * pos;
    r[0] = c.dm;
    r[1] = (float)c.dm_idx;
    r[2] = c.acc;
    r[3] = (float)c.nh;
    r[4] = c.snr;
    r[5] = c.freq;
    nassoc[pos] = c.count_assoc();
  }
  ++pos;
  for (size_t k = 0; k < c.assoc.size(); ++k) rows_from_tree(c.assoc[k], rows, nassoc, cap, pos);
}

void score_rows(std::vector<Candidate>& v, float* ratios, int* flags) {
  for (size_t t = 0; t < v.size(); ++t) {
    ratios[2 * t] = v[t].ddm_count_ratio;
    ratios[2 * t + 1] = v[t].ddm_snr_ratio;
    flags[2 * t] = v[t].is_adjacent;
    flags[2 * t + 1] = v[t].is_physical;
  }
}
}  // namespace

// CandidateScorer::score_all over the top-level candidates of a forest.
extern "C" int ref_score(const float* rows, const int* nassoc, int n, float tsamp, float cfreq, float foff,
                         float bw, float* ratios, int* flags) {
  std::vector<Candidate> v;
  if (forest_from_rows(rows, nassoc, n, v)) return -1;
  CandidateScorer sc(tsamp, cfreq, foff, bw);
  sc.score_all(v);
  score_rows(v, ratios, flags);
  return (int)v.size();
}

// main() after the searches (src/pipeline_multi.cu:406-407,547-558).
extern "C" int ref_finalize(const float* rows, const int* nassoc, int n, float freq_tol, float max_harm,
                            float tsamp, float cfreq, float foff, float bw, int limit, float* out_rows,
                            int* out_nassoc, int cap, float* ratios, int* flags, int* nfinal) {
  std::vector<Candidate> v;
  if (forest_from_rows(rows, nassoc, n, v)) return -1;
  DMDistiller dm_still(freq_tol, true);
  HarmonicDistiller harm_still(freq_tol, max_harm, true, false);
  v = dm_still.distill(v);
  v = harm_still.distill(v);
  CandidateScorer sc(tsamp, cfreq, foff, bw);
  sc.score_all(v);
  int new_size = std::min(limit, (int)v.size());
  v.resize(new_size);
  score_rows(v, ratios, flags);
  *nfinal = (int)v.size();
  int pos = 0;
  for (size_t t = 0; t < v.size(); ++t) rows_from_tree(v[t], out_rows, out_nassoc, cap, pos);
  return pos;
}

extern "C" int ref_harmonic_distill(const float* rows, int n, float tol, float max_harm, int keep, int frac,
                                    float* out, int* nassoc) {
  std::vector<Candidate> v = rows_to_cands(rows, n);
  HarmonicDistiller d(tol, max_harm, keep != 0, frac != 0);
  std::vector<Candidate> r = d.distill(v);
  return cands_to_rows(r, out, nassoc);
}

extern "C" int ref_accel_distill(const float* rows, int n, float tobs, float tol, int keep, float* out,
                                 int* nassoc) {
  std::vector<Candidate> v = rows_to_cands(rows, n);
  AccelerationDistiller d(tobs, tol, keep != 0);
  std::vector<Candidate> r = d.distill(v);
  return cands_to_rows(r, out, nassoc);
}

extern "C" int ref_dm_distill(const float* rows, int n, float tol, int keep, float* out, int* nassoc) {
  std::vector<Candidate> v = rows_to_cands(rows, n);
  DMDistiller d(tol, keep != 0);
  std::vector<Candidate> r = d.distill(v);
  return cands_to_rows(r, out, nassoc);
}
