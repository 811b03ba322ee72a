// Status handling and version of the C ABI.
#include "common.hpp"
#include "cugal_hip.h"

#include <cstdarg>
#include <cstdio>

namespace cg {

namespace {
thread_local char g_error[512] = "";
}

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
    return code;
}

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(CG_ELAUNCH, "%s launch failed: %s", what, hipGetErrorString(e));
    return CG_OK;
}

}  // namespace cg

extern "C" {

const char* cg_last_error(void) { return cg::g_error; }

int cg_version(void) { return 100; }

}  // extern "C"
