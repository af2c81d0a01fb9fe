// Error reporting and version query of the C ABI.
#include <cstdio>

#include "common.h"

namespace sv {

static thread_local char g_err[1024] = "";

int set_error(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}

}  // namespace sv

extern "C" {

const char* sv_last_error(void) { return sv::g_err; }

int sv_version(void) { return 1; }

}  // extern "C"
