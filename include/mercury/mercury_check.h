// C++ error convention of the drop-in headers: a failed C-ABI call throws
// std::runtime_error with the library's message (the reference throws
// std::runtime_error or aborts on the same conditions).
#pragma once

#include <stdexcept>
#include <string>

#include "../mercury_hip.h"

namespace mercury_hip_detail {
inline void check(int rc, const char *what) {
    if (rc != 0) throw std::runtime_error(std::string(what) + ": " + mercury_last_error());
}
}  // namespace mercury_hip_detail
