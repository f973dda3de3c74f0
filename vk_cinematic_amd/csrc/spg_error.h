// Error reporting of the C ABI: a negative SPG_E_* status plus a
// thread-local message (spg_last_error).  Host-only, no HIP.
#pragma once

#include <string>

#include "../../include/spgpu.h"

namespace spg {
int fail(int code, const std::string &msg);
}  // namespace spg

#define CHK(x)                  \
    do                          \
    {                           \
        int _r = (x);           \
        if (_r) return _r;      \
    } while (0)
