// Host-side C ABI plumbing: error reporting, version, device discovery.
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>

#include "sbz_internal.h"

static thread_local char g_err[1024] = "";

void sbz_set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

void sbz_clear_error() { g_err[0] = '\0'; }

extern "C" const char* sbz_last_error(void) { return g_err; }

extern "C" int sbz_version(void) { return 100; }

extern "C" int sbz_device_count(int* count) {
    if (!count) return SBZ_ERR_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return SBZ_OK;
}
