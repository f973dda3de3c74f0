// Thread-local error message of the C ABI (host-only).
#include "spg_error.h"

namespace spg {

static thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

const char *last_error() { return g_err.c_str(); }

}  // namespace spg

extern "C" const char *spg_last_error(void) { return spg::last_error(); }
