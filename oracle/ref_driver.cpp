// ref_driver.cpp - C entry points over the reference's own distillers.
//
// Compiled by oracle/ref.mk against the reference sources (include/transforms/
// distiller.hpp and include/data_types/candidates.hpp, which need only the C++
// standard library) into oracle/_ref/libref.so.  Test infrastructure only: it
// checks the oracle's restatement of the distillers.
#include "transforms/distiller.hpp"

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
}  // namespace

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
