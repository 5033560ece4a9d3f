// postprocess.cpp - what the search program does with the candidates of all
// DM trials after the searches (src/pipeline_multi.cu:545-561):
//   DMDistiller(freq_tol, keep related)
//   -> HarmonicDistiller(freq_tol, max_harm, keep related, no fractional)
//   -> CandidateScorer(tsamp, cfreq, foff, |foff| * nchans)
//   -> keep the first --limit candidates
// and the candidates.peasoup binary file it writes
// (CandidateFileWriter::write_binary, include/utils/output_stats.hpp:261-298).
// Host code only: no device calls.
#include <cstdio>
#include <cstring>
#include <exception>
#include <vector>

#include "peasoup/candidates.hpp"
#include "peasoup/distiller.hpp"
#include "peasoup/scorer.hpp"
#include "search_common.hpp"

namespace psx {
namespace search {

namespace {

void write_tree(const Candidate& c, psx_candidate* out, uint32_t cap, uint64_t& pos) {
  if (pos < cap) out[pos] = psx_candidate{c.dm, c.dm_idx, c.acc, c.nh, c.snr, c.freq, c.count_assoc(), 0};
  ++pos;
  for (const Candidate& a : c.assoc) write_tree(a, out, cap, pos);
}

// Subtree rooted at in[i]; returns the index after it.
uint32_t read_tree(const psx_candidate* in, uint32_t i, Candidate& c) {
  const psx_candidate& r = in[i];
  c = Candidate(r.dm, r.dm_idx, r.acc, r.nh, r.snr, r.freq);
  const uint32_t end = i + 1 + (uint32_t)r.nassoc;
  uint32_t j = i + 1;
  while (j < end) {
    c.assoc.emplace_back();
    j = read_tree(in, j, c.assoc.back());
  }
  return j;
}

// Every subtree must end inside its parent's range and the forest inside n.
bool well_formed(const psx_candidate* in, uint32_t n) {
  std::vector<uint64_t> ends;  // open subtrees' end indices, innermost last
  for (uint32_t i = 0; i < n; ++i) {
    while (!ends.empty() && ends.back() == i) ends.pop_back();
    if (in[i].nassoc < 0) return false;
    const uint64_t e = (uint64_t)i + 1 + (uint64_t)in[i].nassoc;
    if (e > n || (!ends.empty() && e > ends.back())) return false;
    ends.push_back(e);
  }
  return true;
}

}  // namespace

uint64_t forest_size(const std::vector<Candidate>& cands) {
  uint64_t n = 0;
  for (const Candidate& c : cands) n += 1 + (uint64_t)c.count_assoc();
  return n;
}

int flatten(const std::vector<Candidate>& cands, psx_candidate* out, uint32_t cap, uint32_t* count) {
  uint64_t pos = 0;
  for (const Candidate& c : cands) write_tree(c, out, cap, pos);
  if (pos > 0xFFFFFFFFull) return psx::fail(PSX_ERR_OVERFLOW, "more than 2^32 candidate records");
  if (count) *count = (uint32_t)pos;
  if (pos > cap) return psx::fail(PSX_ERR_OVERFLOW, "candidate output holds %u of %llu records", cap,
                                  (unsigned long long)pos);
  return PSX_OK;
}

int unflatten(const psx_candidate* in, uint32_t n, std::vector<Candidate>& cands) {
  PSX_REQUIRE(in || n == 0, "null candidate records");
  PSX_REQUIRE(well_formed(in, n), "malformed candidate forest (nassoc overruns its parent or the input)");
  cands.clear();
  uint32_t i = 0;
  while (i < n) {
    cands.emplace_back();
    i = read_tree(in, i, cands.back());
  }
  return PSX_OK;
}

}  // namespace search
}  // namespace psx

extern "C" int psx_finalize(const psx_candidate* in, uint32_t nin, const psx_final_args* a, psx_candidate* out,
                            uint32_t cap, uint32_t* count, psx_cand_score* scores, uint32_t scap, uint32_t* nfinal) {
  PSX_REQUIRE(a, "psx_finalize: null arguments");
  PSX_REQUIRE(a->limit >= 0, "psx_finalize: limit must be >= 0");
  PSX_REQUIRE(a->cfreq > 0 && a->bw >= 0 && a->bw < 2 * a->cfreq, "psx_finalize: need 0 <= bw < 2 cfreq");
  try {
    std::vector<Candidate> cands;
    int rc = psx::search::unflatten(in, nin, cands);
    if (rc != PSX_OK) return rc;
    DMDistiller dm_still(a->freq_tol, true);
    HarmonicDistiller harm_still(a->freq_tol, a->max_harm, true, false);
    cands = dm_still.distill(cands);
    cands = harm_still.distill(cands);
    CandidateScorer scorer(a->tsamp, a->cfreq, a->foff, a->bw);
    scorer.score_all(cands);
    if ((size_t)a->limit < cands.size()) cands.resize((size_t)a->limit);
    if (nfinal) *nfinal = (uint32_t)cands.size();
    if (scores) {
      for (size_t i = 0; i < cands.size() && i < scap; ++i) {
        const Candidate& c = cands[i];
        scores[i] = psx_cand_score{c.ddm_count_ratio, c.ddm_snr_ratio,     (int32_t)c.is_adjacent,
                                   (int32_t)c.is_physical, c.count_assoc(), (int32_t)c.assoc.size()};
      }
      if (cands.size() > scap)
        return psx::fail(PSX_ERR_OVERFLOW, "score output holds %u of %zu candidates", scap, cands.size());
    }
    rc = psx::search::flatten(cands, out, cap, count);
    return rc == PSX_OK ? psx::ok() : rc;
  } catch (const std::exception& e) {
    return psx::fail(PSX_ERR_ALLOC, "psx_finalize: %s", e.what());
  }
}

extern "C" int psx_write_candidates(const char* path, const psx_candidate* forest, uint32_t n,
                                    const psx_fold* folds, int64_t* offsets, uint32_t ntop_cap, uint32_t* ntop) {
  PSX_REQUIRE(path, "psx_write_candidates: null path");
  PSX_REQUIRE(forest || n == 0, "psx_write_candidates: null records");
  std::vector<Candidate> check;
  int rc = psx::search::unflatten(forest, n, check);
  if (rc != PSX_OK) return rc;
  const uint32_t tops = (uint32_t)check.size();
  if (ntop) *ntop = tops;
  PSX_REQUIRE(!(folds || offsets) || tops <= ntop_cap, "psx_write_candidates: %u candidates, arrays hold %u", tops,
              ntop_cap);
  FILE* fo = std::fopen(path, "wb");
  PSX_REQUIRE(fo, "psx_write_candidates: cannot open %s", path);
  bool ok = true;
  std::vector<CandidatePOD> dets;
  uint32_t i = 0;
  for (uint32_t t = 0; t < tops && ok; ++t) {
    if (offsets) offsets[t] = (int64_t)std::ftell(fo);
    if (folds && (int64_t)folds[t].nbins * folds[t].nints > 0) {
      const psx_fold& f = folds[t];
      const size_t nf = (size_t)f.nbins * (size_t)f.nints;
      ok = ok && f.data && std::fwrite("FOLD", 1, 4, fo) == 4;
      ok = ok && std::fwrite(&f.nbins, sizeof(int32_t), 1, fo) == 1;
      ok = ok && std::fwrite(&f.nints, sizeof(int32_t), 1, fo) == 1;
      ok = ok && std::fwrite(f.data, sizeof(float), nf, fo) == nf;
    }
    const int32_t ndets = 1 + forest[i].nassoc;
    dets.resize((size_t)ndets);
    for (int32_t k = 0; k < ndets; ++k) {
      const psx_candidate& r = forest[i + (uint32_t)k];
      dets[(size_t)k] = CandidatePOD{r.dm, r.dm_idx, r.acc, r.nh, r.snr, r.freq};
    }
    ok = ok && std::fwrite(&ndets, sizeof(int32_t), 1, fo) == 1;
    ok = ok && std::fwrite(dets.data(), sizeof(CandidatePOD), (size_t)ndets, fo) == (size_t)ndets;
    i += (uint32_t)ndets;
  }
  ok = (std::fclose(fo) == 0) && ok;
  PSX_REQUIRE(ok, "psx_write_candidates: write to %s failed", path);
  return psx::ok();
}
