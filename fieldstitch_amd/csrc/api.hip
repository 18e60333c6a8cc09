// C-ABI plumbing of libfsx: thread-local error text and device queries.
#include "fsx_common.h"
#include <string.h>

namespace fsx {

static thread_local char g_err[1024];

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

}  // namespace fsx

extern "C" {

int fsx_version(void) { return 1; }

// Copies the calling thread's last error message (empty if none) into buf; returns its length.
int fsx_last_error(char* buf, size_t len) {
    size_t n = strlen(fsx::g_err);
    if (buf && len) {
        size_t c = n < len - 1 ? n : len - 1;
        memcpy(buf, fsx::g_err, c);
        buf[c] = 0;
    }
    return (int)n;
}

int fsx_device_info(int device, char* name, size_t len, int* n_cu, long long* hbm_bytes) {
    hipDeviceProp_t p;
    FSX_HIP(hipGetDeviceProperties(&p, device));
    if (name && len) {
        snprintf(name, len, "%s (%s)", p.name, p.gcnArchName);
    }
    if (n_cu) *n_cu = p.multiProcessorCount;
    if (hbm_bytes) *hbm_bytes = (long long)p.totalGlobalMem;
    return FSX_OK;
}

}  // extern "C"
