// Shared definitions of libfsx: error state, launch checks and the SIFT workspace layout.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdio.h>
#include <stdarg.h>

#define FSX_OK 0
#define FSX_EINVAL 1
#define FSX_EHIP 2
#define FSX_ECAPACITY 3
#define FSX_ENOMEM 4

namespace fsx {

void set_error(const char* fmt, ...);

// Host -> device upload of a small table on `st` through library-owned pinned staging memory.
// hipMemcpyAsync may read pageable host memory after it returns, so callers never pass their
// own (stack / vector) memory: the bytes are copied into a pinned slot first, and a slot is
// reused only after the copy that last read it has completed (event wait).
int upload_async(void* dst_dev, const void* src, size_t n, hipStream_t st);

// Launch-shape tunables (process-wide; set by the autotuner through fsx_tune_set)
enum TuneKey {
    TUNE_ORIENT_BLOCKS = 0,    // orientation blocks (one wave each) per image
    TUNE_REFINE_BLOCKS = 1,    // refinement blocks per image
    TUNE_FSCREEN_TILE = 2,     // float matcher: targets per LDS tile (64 or 128)
    TUNE_FSCREEN_QB = 3,       // float matcher: blocks of 32 queries per wave (1 or 2)
    TUNE_BLUR_ROWS = 4,        // pyramid blur: output rows per tile, the row strip (8, 16, 24 or 32)
    TUNE_MATCH_BLOCKS = 5,     // guided matcher contraction: 0 = one workgroup per (chunk, pair),
                               // N = N persistent workgroups striding over the items
    TUNE_N_KEYS = 6
};
int tune_get(int key);

#define FSX_HIP(call)                                                                \
    do {                                                                             \
        hipError_t e_ = (call);                                                      \
        if (e_ != hipSuccess) {                                                      \
            ::fsx::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call, hipGetErrorString(e_)); \
            return FSX_EHIP;                                                         \
        }                                                                            \
    } while (0)

#define FSX_LAUNCH_CHECK()                                                           \
    do {                                                                             \
        hipError_t e_ = hipGetLastError();                                           \
        if (e_ != hipSuccess) {                                                      \
            ::fsx::set_error("%s:%d kernel launch: %s", __FILE__, __LINE__, hipGetErrorString(e_)); \
            return FSX_EHIP;                                                         \
        }                                                                            \
    } while (0)

#define FSX_REQUIRE(cond, ...)                                                       \
    do {                                                                             \
        if (!(cond)) {                                                               \
            ::fsx::set_error(__VA_ARGS__);                                           \
            return FSX_EINVAL;                                                       \
        }                                                                            \
    } while (0)

// ------------------------------------------------------------------------------------------
// SIFT constants (gsift.py:13-25, 224)
constexpr int N_LAYERS = 3;
constexpr int BORDER = 5;
constexpr int MAX_INTERP = 5;
constexpr int ORI_BINS = 36;
constexpr int ORI_R_MAX = 16;
constexpr int MAXO = 16;
constexpr int HIST_SIZE = 6 * 6 * 10;    // (d+2) x (d+2) x (bins+2) descriptor histogram

// Octave geometry inside one image's workspace.  Offsets are in float elements.
struct OctGeom {
    int R, C;
    long long g45base;   // gauss layers 4..5: (2, R, C) per octave
    long long gbase;     // gauss layers 1..3: (3, R, C) per octave
    long long g0base;    // gauss layer 0 (octave base): (R, C) per octave
};

// Per-image job record, stored at the head of the image's workspace (device memory).
struct SiftJob {
    const uint8_t* img;
    int H, W, ld;
    int n_oct;
    OctGeom oct[MAXO];
    float* G0;           // sumP: gauss layer 0 of every octave
    float* G123;         // 3 * sumP: gauss layers 1..3
    float* G45;          // 2 * sumP: gauss layers 4..5 (DoG layer l = G[l+1] - G[l], never stored)
    // keypoint pipeline buffers
    uint32_t* cand;      // candidate atlas indices (gauss-atlas numbering, layers 1..3)
    float* cand_der;     // [cand_cap][10] DoG derivatives at each candidate (v, dx, dy, ds, dxx, dyy, dss, dxy, dxs, dys)
    float4* refined;     // 2 x float4 per refined keypoint: (o,l,r,c as int bits), (X0,X1,X2,resp)
    uint32_t* refined_src;   // candidate index of each refined keypoint
    float4* oriented;    // 2 x float4 per oriented keypoint: (x,y,size,angle), (resp, o, l bits, src)
    unsigned long long* keys0;   // sort keys (2 per element) ping
    unsigned long long* keys1;   // pong
    int* counters;       // [0] n_cand, [1] n_refined, [2] n_oriented, [3] overflow flags, [4] n selected
    int cand_cap, kp_cap;
    float contrast_f, edge_f, e1sq_f, pthr;
    int max_kp;
    float* kp_out;       // [7][max_kp] x,y,size,angle,response,octave(int),layer(int)
    int* n_out;          // number of keypoints written (device)
    // feature-batch extension: grid spread of the detected keypoints, then descriptors
    int nfeat;           // keypoints kept per image (features._spread n)
    long long* sel;      // [nfeat] selected keypoint indices in final (response) order
    int* spread_tmp;     // [4 * kp_cap] scratch of the spread kernel
    float* desc_out;     // [nfeat][128] float32 descriptors (may be null)
    uint8_t* u8_out;     // [nfeat][128] RootSIFT uint8 (may be null)
};

struct SiftPlan {
    int H, W, n_oct;
    int R[MAXO], C[MAXO];
    long long sumP, P0;
};

int sift_plan(int H, int W, SiftPlan* p);

}  // namespace fsx
