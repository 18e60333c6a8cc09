// C-ABI plumbing of libfsx: thread-local error text and device queries.
#include "fsx_common.h"
#include <mutex>
#include <string.h>

namespace fsx {

static thread_local char g_err[1024];

// tunable launch shapes: orientation blocks / image, refinement blocks / image, float-screen tile
static int g_tune[TUNE_N_KEYS] = {4096, 256, 64, 1, 16, 0};
int tune_get(int key) { return g_tune[key]; }

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

namespace {
struct Slot {
    void* buf = nullptr;
    size_t cap = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
constexpr int N_SLOTS = 64;
Slot g_slots[N_SLOTS];
int g_next = 0;
std::mutex g_mu;
}  // namespace

int upload_async(void* dst_dev, const void* src, size_t n, hipStream_t st) {
    if (n == 0) return FSX_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    Slot& s = g_slots[g_next];
    g_next = (g_next + 1) % N_SLOTS;
    if (!s.ev) FSX_HIP(hipEventCreateWithFlags(&s.ev, hipEventDisableTiming));
    if (s.pending) {
        FSX_HIP(hipEventSynchronize(s.ev));
        s.pending = false;
    }
    if (s.cap < n) {
        if (s.buf) FSX_HIP(hipHostFree(s.buf));
        s.buf = nullptr;
        size_t cap = 4096;
        while (cap < n) cap <<= 1;
        FSX_HIP(hipHostMalloc(&s.buf, cap, hipHostMallocDefault));
        s.cap = cap;
    }
    memcpy(s.buf, src, n);
    FSX_HIP(hipMemcpyAsync(dst_dev, s.buf, n, hipMemcpyHostToDevice, st));
    FSX_HIP(hipEventRecord(s.ev, st));
    s.pending = true;
    return FSX_OK;
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

int fsx_tune_set(int key, int value) {
    using namespace fsx;
    FSX_REQUIRE(key >= 0 && key < TUNE_N_KEYS, "unknown tuning key %d", key);
    if (key == TUNE_ORIENT_BLOCKS || key == TUNE_REFINE_BLOCKS)
        FSX_REQUIRE(value >= 16 && value <= 65536, "blocks per image must be in [16, 65536] (got %d)", value);
    if (key == TUNE_FSCREEN_TILE) FSX_REQUIRE(value == 64 || value == 128, "tile must be 64 or 128 (got %d)", value);
    if (key == TUNE_FSCREEN_QB) FSX_REQUIRE(value == 1 || value == 2, "query blocks must be 1 or 2 (got %d)", value);
    if (key == TUNE_BLUR_ROWS)
        FSX_REQUIRE(value == 8 || value == 16 || value == 24 || value == 32, "blur rows must be 8, 16, 24 or 32 (got %d)",
                    value);
    if (key == TUNE_MATCH_BLOCKS)
        FSX_REQUIRE(value == 0 || (value >= 256 && value <= 65536),
                    "matcher workgroups must be 0 (one per chunk) or in [256, 65536] (got %d)", value);
    g_tune[key] = value;
    return FSX_OK;
}

int fsx_tune_get(int key, int* value) {
    using namespace fsx;
    FSX_REQUIRE(key >= 0 && key < TUNE_N_KEYS && value, "unknown tuning key %d", key);
    *value = g_tune[key];
    return FSX_OK;
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
