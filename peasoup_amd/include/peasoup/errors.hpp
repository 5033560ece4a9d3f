// errors.hpp - status-to-exception bridge for the C++ transform classes.
// The reference throws std::runtime_error from ErrorChecker after
// synchronising the device on every call (include/utils/exceptions.hpp:13-74);
// these classes throw on a non-zero C-ABI status instead, without a sync.
#pragma once
#include <stdexcept>
#include <string>

#include "psx.h"

namespace peasoup {

inline void check(int rc, const char* what) {
  if (rc != PSX_OK) throw std::runtime_error(std::string(what) + ": " + psx_last_error_string());
}

inline void throw_error(const std::string& msg) { throw std::runtime_error(msg); }

}  // namespace peasoup
