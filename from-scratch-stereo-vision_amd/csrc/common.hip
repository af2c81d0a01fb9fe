// Error reporting, version query and tuning knobs of the C ABI.
#include <atomic>
#include <cstdio>

#include "common.h"

namespace sv {

static thread_local char g_err[1024] = "";
static std::atomic<int> g_tune[SV_TUNE_COUNT];

int tune(int knob) { return knob >= 0 && knob < SV_TUNE_COUNT ? g_tune[knob].load() : 0; }

int cu_count() {
    static std::atomic<int> cached[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0) return 256;
    if (dev < 64) {
        const int c = cached[dev].load();
        if (c > 0) return c;
    }
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || c <= 0)
        c = 256;
    if (dev < 64) cached[dev].store(c);
    return c;
}

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

// Allowed values per knob (0 always allowed: back to the heuristic).
int sv_tune_set(int knob, int value) {
    static const int allowed[SV_TUNE_COUNT][7] = {
        {2, 4, 8, -1, -1, -1, -1}, {8, 16, -1, -1, -1, -1, -1}, {1, 4, -1, -1, -1, -1, -1},
        {16, 32, 64, -1, -1, -1, -1}, {1, 2, -1, -1, -1, -1, -1}, {1, 2, -1, -1, -1, -1, -1},
        {1, 2, 3, 4, 5, 6, 7}, {1, 2, 4, -1, -1, -1, -1}, {1, 2, -1, -1, -1, -1, -1},
        {1, 2, -1, -1, -1, -1, -1}, {4, 6, 12, -1, -1, -1, -1}, {1, 2, -1, -1, -1, -1, -1},
        {1, 2, 3, -1, -1, -1, -1}};
    if (knob < 0 || knob >= SV_TUNE_COUNT)
        return sv::set_error(SV_EINVAL, "tuning knob %d does not exist", knob);
    bool ok = value == 0;
    for (int i = 0; i < 7; ++i) ok = ok || allowed[knob][i] == value;
    if (!ok) return sv::set_error(SV_EINVAL, "value %d is not allowed for tuning knob %d", value, knob);
    sv::g_tune[knob].store(value);
    return SV_OK;
}

int sv_tune_get(int knob, int* value) {
    if (knob < 0 || knob >= SV_TUNE_COUNT || !value)
        return sv::set_error(SV_EINVAL, "tuning knob %d does not exist", knob);
    *value = sv::g_tune[knob].load();
    return SV_OK;
}

}  // extern "C"
