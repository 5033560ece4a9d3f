// search_common.hpp - helpers shared by the Worker-level search drivers.
#pragma once
#include <vector>

#include "common.hpp"
#include "peasoup/candidates.hpp"

namespace psx {
namespace search {

// Writes candidates as psx_candidate records; PSX_ERR_OVERFLOW past cap.
int flatten(const std::vector<Candidate>& cands, psx_candidate* out, uint32_t cap, uint32_t* count);
int validate_args(const psx_search_args* a);

}  // namespace search
}  // namespace psx
