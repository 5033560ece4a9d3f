// search_common.hpp - helpers shared by the Worker-level search drivers and
// the global post-processing.
#pragma once
#include <vector>

#include "common.hpp"
#include "peasoup/candidates.hpp"

namespace psx {

namespace search {

// Candidate lists cross the C ABI as a "forest": each top-level candidate is
// followed by its associated candidates in pre-order (the order of
// Candidate::collect_candidates, include/data_types/candidates.hpp:78-84),
// and every record's nassoc is the number of records of its own subtree
// after it (Candidate::count_assoc).
//
// Writes cands as a forest; *count = records in it; PSX_ERR_OVERFLOW past cap.
int flatten(const std::vector<Candidate>& cands, psx_candidate* out, uint32_t cap, uint32_t* count);
// Number of records flatten writes for cands.
uint64_t forest_size(const std::vector<Candidate>& cands);
// Rebuilds the trees of a forest; PSX_ERR_ARG if a subtree overruns n.
int unflatten(const psx_candidate* in, uint32_t n, std::vector<Candidate>& cands);
int validate_args(const psx_search_args* a);

}  // namespace search
}  // namespace psx
