// Descriptor matcher for gfx950: the prior-guided matcher (app/planesite/matching.py:113-152
// _guided) and the brute-force ratio matcher (matching.py:86-100 _ratio_match_bf), each with
// the one-query-per-target rule (matching.py:103-110 _unique), batched over many image pairs.
//
// Match set definition (every path here returns exactly this list): candidates of query q =
// the B keypoints strictly inside the prior radius around q's predicted position (squared
// float64 distance), optionally only the kc spatially nearest (ties: lower index), or all of
// B for brute force; d = sum_k (a_k - b_k)^2; best = argmin d (ties: lower index), b2 = the
// second smallest d (+inf with one candidate); accept iff sqrt(b1) < ratio*sqrt(b2) in
// float64; per target keep the smallest b1, ties to the lower query; output sorted by target.
// uint8 RootSIFT descriptors make every d an exact integer < 2^24, so the float32 forms and
// the BF16-input MFMA form (uint8 is exact in bf16, sums exact in float32) are exact.
//
// Two forms of the candidate contraction:
//   plain: one thread per query, FP32 gather + dot products on VALU (the baseline form);
//   mfma:  32 spatially neighbouring queries per wave (each keypoint set is bucketed by grid
//          cell in tile-major cell order, which the spatial grid needs anyway), the union of
//          their candidate cells contracted on MFMA (32x32x16 BF16, FP32 accumulation) with
//          the prior-window mask and the top-2 selection fused into the epilogue.
#include "fsx_common.h"
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

namespace fsx {

struct MImage {
    long long off;      // first keypoint in the concatenated arrays
    int n;              // keypoints
    int W, H;           // image size (inside test on the B side)
    int gw, gh;         // spatial grid (cells of `radius` px covering [-2r, W+2r) x [-2r, H+2r))
    int tw4;            // tiles of 4x4 cells per tile row; cells are numbered tile-major
    int ncell;          // cell slots (whole tiles)
    long long cell_off; // first cell of this image in the cell arrays
};

struct MPair {
    int a, b;           // image indices: queries from a, targets from b
    int dense;          // 1: brute force over all of b (no prior)
    int pad;
    double Hm[9];       // prior mapping a px -> b px (cv2.perspectiveTransform semantics)
    long long q_off;    // first global query (prefix of n_a)
    long long t_off;    // first per-target slot (prefix of n_b)
    long long o_off;    // first output slot (prefix of min(n_a, n_b))
    long long c_off;    // first MFMA chunk (prefix of ceil(n_a / 32))
};

struct MatchBatch {
    const uint8_t* desc;          // [Ntot][128] RootSIFT uint8
    const double* xy;             // [Ntot][2] positions (undistorted px)
    const MImage* img;
    const MPair* pair;
    int n_img, n_pair;
    long long n_query, n_kp, n_cells, n_targets;
    double radius, r2, ratio;
    int kc;
    int pad;
    // scratch
    unsigned* norms;              // [Ntot] sum of squares
    int* cell_start;              // [n_cells + n_img]
    int* cell_cur;                // [n_cells]
    int* cell_idx;                // [Ntot] local keypoint indices sorted by cell
    double* pred;                 // [n_query][2]
    int* qcell;                   // [n_query] cell of the prediction in b's grid, -1 = not inside
    int* best;                    // [n_query] accepted best target (local), -1 = none
    double* b1;                   // [n_query]
    double* kth;                  // [n_query][2] kc-th nearest (d2, idx) threshold, kc > 0
    unsigned long long* tkey;     // [n_targets] min b1 bits per target
    int* tq;                      // [n_targets] min query among equal b1
    int* counters;                // [0..3] flags
    long long n_chunks;           // sum over pairs of ceil(n_a / 32)
    int* out_q;                   // [Σ min(n_a, n_b)]
    int* out_t;
    int* out_n;                   // [n_pair]
};

// ------------------------------------------------------------------------------------------
__device__ __forceinline__ int find_pair(const MatchBatch& B, long long g) {
    int lo = 0, hi = B.n_pair - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (B.pair[mid].q_off <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int find_image(const MatchBatch& B, long long k) {
    int lo = 0, hi = B.n_img - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (B.img[mid].off <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// tile-major cell numbering: 4x4-cell tiles row by row, cells row by row inside a tile, so a
// run of consecutive cells (and of the keypoints bucketed into them) is spatially compact
__device__ __forceinline__ int cell_id(const MImage& im, int cx, int cy) {
    return (((cy >> 2) * im.tw4 + (cx >> 2)) << 4) | ((cy & 3) << 2) | (cx & 3);
}

__device__ __forceinline__ void cell_xy(const MImage& im, int id, int& cx, int& cy) {
    const int tile = id >> 4;
    cy = (tile / im.tw4) * 4 + ((id >> 2) & 3);
    cx = (tile % im.tw4) * 4 + (id & 3);
}

__device__ __forceinline__ void cell_of(const MatchBatch& B, const MImage& im, double x, double y, int& cx, int& cy) {
    cx = (int)floor((x + 2.0 * B.radius) / B.radius);
    cy = (int)floor((y + 2.0 * B.radius) / B.radius);
    cx = cx < 0 ? 0 : (cx >= im.gw ? im.gw - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= im.gh ? im.gh - 1 : cy);
}

__global__ void k_m_reset(MatchBatch B) {
    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x, stride = (long long)gridDim.x * 256;
    for (long long i = tid; i < B.n_targets; i += stride) {
        B.tkey[i] = ~0ull;
        B.tq[i] = 0x7fffffff;
    }
    for (long long i = tid; i < B.n_cells; i += stride) B.cell_cur[i] = 0;
    for (long long i = tid; i < B.n_pair; i += stride) B.out_n[i] = 0;
    if (tid < 4) B.counters[tid] = 0;
}

// sum of squares of every descriptor (exact integer for uint8) and its grid cell count
__global__ void k_m_norms_count(MatchBatch B) {
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= B.n_kp) return;
    const uint4* d = (const uint4*)(B.desc + k * 128);
    unsigned s = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        uint4 v = d[i];
        s = __builtin_amdgcn_udot4(v.x, v.x, s, false);
        s = __builtin_amdgcn_udot4(v.y, v.y, s, false);
        s = __builtin_amdgcn_udot4(v.z, v.z, s, false);
        s = __builtin_amdgcn_udot4(v.w, v.w, s, false);
    }
    B.norms[k] = s;
    const int ii = find_image(B, k);
    const MImage& im = B.img[ii];
    int cx, cy;
    cell_of(B, im, B.xy[2 * k], B.xy[2 * k + 1], cx, cy);
    atomicAdd(&B.cell_cur[im.cell_off + cell_id(im, cx, cy)], 1);
}

// one block per image: exclusive scan of the cell counts -> cell_start; cursors reset
__global__ __launch_bounds__(256) void k_m_grid_scan(MatchBatch B) {
    __shared__ int s_w[4];
    const MImage& im = B.img[blockIdx.x];
    const int nc = im.ncell;
    int* start = B.cell_start + im.cell_off + blockIdx.x;   // nc + 1 entries per image
    int* cnt = B.cell_cur + im.cell_off;
    int carry = 0;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int base = 0; base < nc; base += 256) {
        const int i = base + tid;
        const int v = i < nc ? cnt[i] : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int pre = 0;
        for (int w2 = 0; w2 < wv; w2++) pre += s_w[w2];
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (i < nc) {
            start[i] = carry + pre + x - v;
            cnt[i] = 0;
        }
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) start[nc] = carry;
}

__global__ void k_m_grid_fill(MatchBatch B) {
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= B.n_kp) return;
    const int ii = find_image(B, k);
    const MImage& im = B.img[ii];
    int cx, cy;
    cell_of(B, im, B.xy[2 * k], B.xy[2 * k + 1], cx, cy);
    const long long c = cell_id(im, cx, cy);
    const int* start = B.cell_start + im.cell_off + ii;
    const int slot = start[c] + atomicAdd(&B.cell_cur[im.cell_off + c], 1);
    B.cell_idx[im.off + slot] = (int)(k - im.off);
}

// prediction (cv2.perspectiveTransform: w = (h6 x + h7 y) + h8, |w| > FLT_EPSILON ? v * (1/w)
// : 0) and the strict inside test of matching.py:119-120
__global__ void k_m_predict(MatchBatch B) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= B.n_query) return;
    const int p = find_pair(B, g);
    const MPair& P = B.pair[p];
    const MImage& ia = B.img[P.a];
    const MImage& ib = B.img[P.b];
    const long long q = g - P.q_off;
    const double x = B.xy[2 * (ia.off + q)], y = B.xy[2 * (ia.off + q) + 1];
    double px, py;
    if (P.dense) {
        px = 0.0;
        py = 0.0;
    } else {
        const double w = (x * P.Hm[6] + y * P.Hm[7]) + P.Hm[8];
        if (fabs(w) > 1.1920928955078125e-07) {
            const double iw = 1.0 / w;
            px = ((x * P.Hm[0] + y * P.Hm[1]) + P.Hm[2]) * iw;
            py = ((x * P.Hm[3] + y * P.Hm[4]) + P.Hm[5]) * iw;
        } else {
            px = 0.0;
            py = 0.0;
        }
    }
    B.pred[2 * g] = px;
    B.pred[2 * g + 1] = py;
    int cell = -1;
    if (P.dense) {
        cell = (ib.n >= 2 && ia.n >= 2) ? 0 : -1;
    } else if (px > -B.radius && px < ib.W + B.radius && py > -B.radius && py < ib.H + B.radius && ib.n > 0) {
        int cx, cy;
        cell_of(B, ib, px, py, cx, cy);
        cell = cell_id(ib, cx, cy);
    }
    B.qcell[g] = cell;
    B.best[g] = -1;
}

// spatial (d2, idx) key comparison for the kc-nearest rule
__device__ __forceinline__ bool sp_less(double d, int i, double d2, int i2) { return d < d2 || (d == d2 && i < i2); }

// kc-th nearest spatial threshold of every query (kc > 0 only): the candidate set is
// {in radius and (d2, idx) <= threshold}
__global__ void k_m_knn(MatchBatch B) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= B.n_query) return;
    const int cell = B.qcell[g];
    double td = B.r2;
    int ti = 0x7fffffff;
    const int p = find_pair(B, g);
    const MPair& P = B.pair[p];
    if (cell >= 0 && !P.dense) {
        constexpr int KMAX = 64;
        double ld[KMAX];
        int li[KMAX];
        int nl = 0;
        const int kc = B.kc < KMAX ? B.kc : KMAX;
        const MImage& ib = B.img[P.b];
        const double px = B.pred[2 * g], py = B.pred[2 * g + 1];
        int cx0, cy0;
        cell_xy(ib, cell, cx0, cy0);
        const int* start = B.cell_start + ib.cell_off + P.b;
        for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
            for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
                const int c = cell_id(ib, cx, cy);
                for (int s = start[c]; s < start[c + 1]; s++) {
                    const int t = B.cell_idx[ib.off + s];
                    const double dx = B.xy[2 * (ib.off + t)] - px, dy = B.xy[2 * (ib.off + t) + 1] - py;
                    const double d = dx * dx + dy * dy;
                    if (!(d < B.r2)) continue;
                    if (nl == kc && !sp_less(d, t, ld[nl - 1], li[nl - 1])) continue;
                    int j = nl < kc ? nl++ : kc - 1;
                    while (j > 0 && sp_less(d, t, ld[j - 1], li[j - 1])) {
                        ld[j] = ld[j - 1];
                        li[j] = li[j - 1];
                        j--;
                    }
                    ld[j] = d;
                    li[j] = t;
                }
            }
        if (nl == kc) {
            td = ld[kc - 1];
            ti = li[kc - 1];
        }
    }
    B.kth[2 * g] = td;
    B.kth[2 * g + 1] = __longlong_as_double((long long)ti);
}

__device__ __forceinline__ bool in_candidates(const MatchBatch& B, long long g, double d, int t) {
    if (!(d < B.r2)) return false;
    if (B.kc <= 0) return true;
    const double td = B.kth[2 * g];
    const int ti = (int)__double_as_longlong(B.kth[2 * g + 1]);
    return d < td || (d == td && t <= ti);
}

// top-2 by (value, index): v1/i1 best, v2 second smallest value
__device__ __forceinline__ void top2_push(float d, int t, float& v1, int& i1, float& v2) {
    if (d < v1 || (d == v1 && t < i1)) {
        v2 = fminf(v2, v1);
        v1 = d;
        i1 = t;
    } else {
        v2 = fminf(v2, d);
    }
}

// ratio test in float64 + the first phase of the per-target minimum
__device__ __forceinline__ void accept_query(const MatchBatch& B, long long g, const MPair& P, float v1, int i1, float v2) {
    if (i1 < 0) return;
    const double b1 = (double)v1, b2 = (double)v2;
    if (sqrt(b1) < B.ratio * sqrt(b2)) {
        B.best[g] = i1;
        B.b1[g] = b1;
        atomicMin(&B.tkey[P.t_off + i1], (unsigned long long)__double_as_longlong(b1));
    }
}

// the plain FP32 GPU form: one thread per query; the query descriptor stays in registers as
// float32, each candidate descriptor is gathered from global memory (16-byte loads) and the
// dot product runs on VALU in FP32: d = na + nb - 2 a.b (exact for uint8 inputs)
__device__ __forceinline__ float u8dot_f32(const float* qa, const uint4* cb) {
    float dot = 0.f;
#pragma unroll
    for (int i = 0; i < 8; i++) {
        const uint4 v = cb[i];
        const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int h = 0; h < 4; h++)
#pragma unroll
            for (int e = 0; e < 4; e++) dot += qa[16 * i + 4 * h + e] * (float)((w[h] >> (8 * e)) & 0xffu);
    }
    return dot;
}

__global__ __launch_bounds__(256) void k_m_plain(MatchBatch B) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= B.n_query) return;
    const int cell = B.qcell[g];
    if (cell < 0) return;
    const int p = find_pair(B, g);
    const MPair& P = B.pair[p];
    const MImage& ia = B.img[P.a];
    const MImage& ib = B.img[P.b];
    const long long ka = ia.off + (g - P.q_off);
    float qa[128];
    {
        const uint4* q4 = (const uint4*)(B.desc + ka * 128);
#pragma unroll
        for (int i = 0; i < 8; i++) {
            const uint4 v = q4[i];
            const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
            for (int h = 0; h < 4; h++)
#pragma unroll
                for (int e = 0; e < 4; e++) qa[16 * i + 4 * h + e] = (float)((w[h] >> (8 * e)) & 0xffu);
        }
    }
    const float na = (float)B.norms[ka];
    float v1 = INFINITY, v2 = INFINITY;
    int i1 = -1;
    auto eval = [&](int t) {
        const float dot = u8dot_f32(qa, (const uint4*)(B.desc + (ib.off + t) * 128));
        const float d = fmaxf((na + (float)B.norms[ib.off + t]) - 2.0f * dot, 0.0f);
        top2_push(d, t, v1, i1, v2);
    };
    if (P.dense) {
        for (int t = 0; t < ib.n; t++) eval(t);
    } else {
        const double px = B.pred[2 * g], py = B.pred[2 * g + 1];
        int cx0, cy0;
        cell_xy(ib, cell, cx0, cy0);
        const int* start = B.cell_start + ib.cell_off + P.b;
        for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
            for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
                const int c = cell_id(ib, cx, cy);
                for (int s = start[c]; s < start[c + 1]; s++) {
                    const int t = B.cell_idx[ib.off + s];
                    const double dx = B.xy[2 * (ib.off + t)] - px, dy = B.xy[2 * (ib.off + t) + 1] - py;
                    if (in_candidates(B, g, dx * dx + dy * dy, t)) eval(t);
                }
            }
    }
    accept_query(B, g, P, v1, i1, v2);
}

// second phase of the per-target minimum: among queries with the minimal b1, the lowest q
__global__ void k_m_unique2(MatchBatch B) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= B.n_query) return;
    const int t = B.best[g];
    if (t < 0) return;
    const int p = find_pair(B, g);
    const MPair& P = B.pair[p];
    if ((unsigned long long)__double_as_longlong(B.b1[g]) == B.tkey[P.t_off + t])
        atomicMin(&B.tq[P.t_off + t], (int)(g - P.q_off));
}

// one block per pair: winners in target order -> (q, t) lists
__global__ __launch_bounds__(256) void k_m_compact(MatchBatch B) {
    __shared__ int s_w[4];
    const MPair& P = B.pair[blockIdx.x];
    const int nb = B.img[P.b].n;
    const int* tq = B.tq + P.t_off;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int carry = 0;
    for (int base = 0; base < nb; base += 256) {
        const int t = base + tid;
        const int v = (t < nb && tq[t] != 0x7fffffff) ? 1 : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int pre = 0;
        for (int w2 = 0; w2 < wv; w2++) pre += s_w[w2];
        const int tot = s_w[0] + s_w[1] + s_w[2] + s_w[3];
        if (v) {
            const long long o = P.o_off + carry + pre + x - 1;
            B.out_q[o] = tq[t];
            B.out_t[o] = t;
        }
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) B.out_n[blockIdx.x] = carry;
}

// ------------------------------------------------------------------------------------------
// MFMA form.  A chunk = 32 consecutive keypoints of the query image in its cell bucketing
// (tile-major cells), so the chunk's predictions land in a compact region of B (the prior is a
// smooth mapping).  One wave per chunk: the union of the queries' 3x3 candidate cells is
// listed once, its candidates are contracted against the 32 queries on MFMA, 32 candidates per
// 32x32x16 BF16 tile (uint8 -> bf16 is exact, FP32 sums of uint8 products are exact), and
// the epilogue applies the prior-window mask and the top-2 update.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int MF_CAP = 512;     // candidates per pass staged in LDS
constexpr int MF_MAXC = 256;    // listed candidate cells per chunk

__device__ __forceinline__ bf16x8_t u8x8_to_bf16(uint2 v) {
    bf16x8_t r;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const unsigned byte = ((e < 4 ? v.x : v.y) >> (8 * (e & 3))) & 0xffu;
        const float f = (float)byte;
        r[e] = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(f) >> 16));
    }
    return r;
}

__device__ __forceinline__ int find_chunk_pair(const MatchBatch& B, long long ch) {
    int lo = 0, hi = B.n_pair - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (B.pair[mid].c_off <= ch) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int wave_excl_scan(int v, int lane, int& total) {
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        const int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    total = __shfl(x, 63);
    return x - v;
}

// per-query evaluation of the 3x3 candidate cells (chunks whose cells do not fit the window)
__device__ void query_direct(const MatchBatch& B, const MPair& P, const MImage& ia, const MImage& ib, int gq, int cell,
                             double qpx, double qpy, float qn, float& v1, int& i1, float& v2) {
    v1 = INFINITY;
    v2 = INFINITY;
    i1 = -1;
    const long long ka = ia.off + (gq - P.q_off);
    const uint8_t* da = B.desc + ka * 128;
    int cx0, cy0;
    cell_xy(ib, cell, cx0, cy0);
    const int* cs = B.cell_start + ib.cell_off + P.b;
    for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
        for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
            const int cc = cell_id(ib, cx, cy);
            for (int s = cs[cc]; s < cs[cc + 1]; s++) {
                const int t = B.cell_idx[ib.off + s];
                const double dx = B.xy[2 * (ib.off + t)] - qpx, dy = B.xy[2 * (ib.off + t) + 1] - qpy;
                if (!in_candidates(B, gq, dx * dx + dy * dy, t)) continue;
                const unsigned* a4 = (const unsigned*)da;
                const unsigned* b4 = (const unsigned*)(B.desc + (ib.off + t) * 128);
                unsigned dot = 0;
                for (int k = 0; k < 32; k++) dot = __builtin_amdgcn_udot4(a4[k], b4[k], dot, false);
                const float d = fmaxf((qn + (float)B.norms[ib.off + t]) - 2.0f * (float)dot, 0.0f);
                top2_push(d, t, v1, i1, v2);
            }
        }
}

__global__ __launch_bounds__(64) void k_m_mfma(MatchBatch B) {
    __shared__ int s_cand[MF_CAP];
    __shared__ double s_cx[MF_CAP], s_cy[MF_CAP];
    __shared__ float s_cn[MF_CAP];
    __shared__ unsigned s_mark[32];
    __shared__ int s_cells[MF_MAXC], s_coff[MF_MAXC + 1];
    __shared__ int s_q[32], s_cell[32];
    __shared__ double s_px[32], s_py[32];
    __shared__ float s_v1[64], s_v2[64];
    __shared__ int s_i1[64];
    const int lane = threadIdx.x;
    for (long long ch = blockIdx.x; ch < B.n_chunks; ch += gridDim.x) {
        const int p = find_chunk_pair(B, ch);
        const MPair& P = B.pair[p];
        const MImage& ia = B.img[P.a];
        const MImage& ib = B.img[P.b];
        const int k0 = (int)(ch - P.c_off) * 32;
        if (lane < 32) {
            int g = -1;
            const int kk = k0 + lane;
            if (kk < ia.n) {
                const long long g0 = P.q_off + B.cell_idx[ia.off + kk];
                if (B.qcell[g0] >= 0) g = (int)g0;
            }
            s_q[lane] = g;
            s_cell[lane] = g >= 0 ? B.qcell[g] : -1;
            s_px[lane] = g >= 0 ? B.pred[2 * (long long)g] : 0.0;
            s_py[lane] = g >= 0 ? B.pred[2 * (long long)g + 1] : 0.0;
            s_mark[lane] = 0u;
        }
        __syncthreads();
        const int qc = lane & 31;
        const int gq = s_q[qc];
        if (__ballot(gq >= 0) == 0ull) {       // wave-uniform: nothing to match in this chunk
            __syncthreads();
            continue;
        }
        // window origin: min cell coordinates of the valid queries, minus one
        int cxq = 1 << 30, cyq = 1 << 30;
        if (gq >= 0) cell_xy(ib, s_cell[qc], cxq, cyq);
        int cx_lo = cxq, cy_lo = cyq;
        for (int off = 32; off >= 1; off >>= 1) {
            cx_lo = min(cx_lo, __shfl_xor(cx_lo, off));
            cy_lo = min(cy_lo, __shfl_xor(cy_lo, off));
        }
        cx_lo = max(cx_lo - 1, 0);
        cy_lo = max(cy_lo - 1, 0);
        bool overflow = false;
        if (!P.dense && lane < 32 && gq >= 0) {
            const int* cs = B.cell_start + ib.cell_off + P.b;
            for (int cy = max(cyq - 1, 0); cy <= min(cyq + 1, ib.gh - 1); cy++)
                for (int cx = max(cxq - 1, 0); cx <= min(cxq + 1, ib.gw - 1); cx++) {
                    const int bx = cx - cx_lo, by = cy - cy_lo;
                    const int cc = cell_id(ib, cx, cy);
                    if (bx < 32 && by < 32 && cs[cc + 1] - cs[cc] <= MF_CAP) atomicOr(&s_mark[by], 1u << bx);
                    else overflow = true;
                }
        }
        bool any_over = __ballot(overflow) != 0ull;
        __syncthreads();
        // list the marked cells (window-row order) and their candidate offsets
        int total = 0;
        int ncell = 0;
        if (P.dense) {
            total = ib.n;
        } else if (!any_over) {
            const unsigned word = lane < 32 ? s_mark[lane] : 0u;
            const int wo = wave_excl_scan(__popc(word), lane, ncell);
            if (ncell > MF_MAXC) {
                any_over = true;
            } else {
                int k = wo;
                for (unsigned w = word; w; w &= w - 1) {
                    const int bx = __ffs(w) - 1;
                    s_cells[k++] = cell_id(ib, cx_lo + bx, cy_lo + lane);
                }
                __syncthreads();
                const int* cs = B.cell_start + ib.cell_off + P.b;
                int carry = 0;
                for (int c0 = 0; c0 < ncell; c0 += 64) {
                    const int c = c0 + lane;
                    const int n = c < ncell ? cs[s_cells[c] + 1] - cs[s_cells[c]] : 0;
                    int tot;
                    const int ex = wave_excl_scan(n, lane, tot);
                    if (c < ncell) s_coff[c] = carry + ex;
                    carry += tot;
                }
                if (lane == 0) s_coff[ncell] = carry;
                total = carry;
                __syncthreads();
            }
        }
        // state of this lane's query column over the candidate passes
        float v1 = INFINITY, v2 = INFINITY;
        int i1 = -1;
        bf16x8_t qop[8];
        float qn = 0.f;
        const double qpx = gq >= 0 ? s_px[qc] : 0.0, qpy = gq >= 0 ? s_py[qc] : 0.0;
        if (gq >= 0) {
            const long long ka = ia.off + (gq - P.q_off);
            const uint2* qa = (const uint2*)(B.desc + ka * 128);
#pragma unroll
            for (int s = 0; s < 8; s++) qop[s] = u8x8_to_bf16(qa[2 * s + (lane >> 5)]);
            qn = (float)B.norms[ka];
        } else {
#pragma unroll
            for (int s = 0; s < 8; s++) qop[s] = u8x8_to_bf16(make_uint2(0, 0));
        }
        if (!any_over) {
            const int* cs = B.cell_start + ib.cell_off + P.b;
            for (int base = 0; base < total; base += MF_CAP) {
                const int np = min(MF_CAP, total - base);
                for (int j = lane; j < np; j += 64) {
                    int t;
                    if (P.dense) {
                        t = base + j;
                    } else {
                        // listed cell holding candidate position base + j: the last cell whose
                        // offset is <= pos (binary search over s_coff[0..ncell))
                        const int pos = base + j;
                        int lo = 0, hi = ncell - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (s_coff[mid] <= pos) lo = mid;
                            else hi = mid - 1;
                        }
                        t = B.cell_idx[ib.off + cs[s_cells[lo]] + (pos - s_coff[lo])];
                        s_cx[j] = B.xy[2 * (ib.off + t)];
                        s_cy[j] = B.xy[2 * (ib.off + t) + 1];
                    }
                    s_cand[j] = t;
                    s_cn[j] = (float)B.norms[ib.off + t];
                }
                __syncthreads();
                for (int t0 = 0; t0 < np; t0 += 32) {
                    const int tr = t0 + (lane & 31);
                    const int tc = tr < np ? s_cand[tr] : -1;
                    const uint2* cb = (const uint2*)(B.desc + (ib.off + (tc >= 0 ? tc : 0)) * 128);
                    f32x16_t acc = {};
#pragma unroll
                    for (int s = 0; s < 8; s++) {
                        const bf16x8_t a = u8x8_to_bf16(tc >= 0 ? cb[2 * s + (lane >> 5)] : make_uint2(0, 0));
                        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, qop[s], acc, 0, 0, 0);
                    }
                    if (gq >= 0) {
#pragma unroll
                        for (int i = 0; i < 4; i++)
#pragma unroll
                            for (int j = 0; j < 4; j++) {
                                const int r = t0 + i * 8 + (lane >> 5) * 4 + j;
                                if (r >= np) continue;
                                const int t = s_cand[r];
                                if (!P.dense) {
                                    const double dx = s_cx[r] - qpx, dy = s_cy[r] - qpy;
                                    if (!in_candidates(B, gq, dx * dx + dy * dy, t)) continue;
                                }
                                const float d = fmaxf((qn + s_cn[r]) - 2.0f * acc[i * 4 + j], 0.0f);
                                top2_push(d, t, v1, i1, v2);
                            }
                    }
                }
                __syncthreads();
            }
        }
        // merge the two half-waves of each query column
        s_v1[lane] = v1;
        s_v2[lane] = v2;
        s_i1[lane] = i1;
        __syncthreads();
        if (lane < 32 && gq >= 0) {
            if (any_over) {
                query_direct(B, P, ia, ib, gq, s_cell[lane], qpx, qpy, qn, v1, i1, v2);
            } else {
                const float w1 = s_v1[lane + 32], w2 = s_v2[lane + 32];
                const int j1 = s_i1[lane + 32];
                if (j1 >= 0) top2_push(w1, j1, v1, i1, v2);
                v2 = fminf(v2, w2);
            }
            accept_query(B, gq, P, v1, i1, v2);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Host side

static inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

struct MatchPlanHost {
    std::vector<MImage> img;
    std::vector<MPair> pair;
    long long n_kp = 0, n_query = 0, n_cells = 0, n_targets = 0, n_out = 0, n_chunks = 0;
    size_t off_img, off_pair, off_norms, off_cstart, off_ccur, off_cidx, off_pred, off_qcell, off_best, off_b1,
        off_kth, off_tkey, off_tq, off_cnt, off_outq, off_outt, off_outn, off_batch, total;
};

static int match_plan(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                      const int* pb, const int* dense, const double* Hm, double radius, MatchPlanHost& M) {
    FSX_REQUIRE(n_img >= 1 && n_pair >= 1, "need images and pairs");
    FSX_REQUIRE(radius > 0, "radius must be > 0");
    M.img.resize(n_img);
    long long off = 0, coff = 0;
    for (int i = 0; i < n_img; i++) {
        FSX_REQUIRE(n_kp[i] >= 0 && W[i] > 0 && H[i] > 0, "bad image %d", i);
        MImage& m = M.img[i];
        m.off = off;
        m.n = n_kp[i];
        m.W = W[i];
        m.H = H[i];
        m.gw = (int)ceil((W[i] + 4.0 * radius) / radius);
        m.gh = (int)ceil((H[i] + 4.0 * radius) / radius);
        m.tw4 = (m.gw + 3) / 4;
        m.ncell = m.tw4 * ((m.gh + 3) / 4) * 16;
        m.cell_off = coff;
        off += n_kp[i];
        coff += m.ncell;
    }
    M.n_kp = off;
    M.n_cells = coff;
    M.pair.resize(n_pair);
    long long qo = 0, to = 0, oo = 0, chunks = 0;
    for (int p = 0; p < n_pair; p++) {
        FSX_REQUIRE(pa[p] >= 0 && pa[p] < n_img && pb[p] >= 0 && pb[p] < n_img, "bad pair %d", p);
        MPair& P = M.pair[p];
        memset(&P, 0, sizeof(P));
        P.a = pa[p];
        P.b = pb[p];
        P.dense = dense ? dense[p] : 0;
        for (int k = 0; k < 9; k++) P.Hm[k] = Hm ? Hm[9 * p + k] : (k % 4 == 0 ? 1.0 : 0.0);
        P.q_off = qo;
        P.t_off = to;
        P.o_off = oo;
        P.c_off = chunks;
        qo += M.img[P.a].n;
        to += M.img[P.b].n;
        oo += std::min(M.img[P.a].n, M.img[P.b].n);
        chunks += (M.img[P.a].n + 31) / 32;
    }
    M.n_query = qo;
    M.n_targets = to;
    M.n_out = oo;
    M.n_chunks = chunks;
    size_t o = 0;
    M.off_batch = o; o = a256(o + sizeof(MatchBatch));
    M.off_img = o; o = a256(o + sizeof(MImage) * n_img);
    M.off_pair = o; o = a256(o + sizeof(MPair) * n_pair);
    M.off_norms = o; o = a256(o + 4 * (size_t)std::max(M.n_kp, 1LL));
    M.off_cstart = o; o = a256(o + 4 * (size_t)(M.n_cells + n_img));
    M.off_ccur = o; o = a256(o + 4 * (size_t)M.n_cells);
    M.off_cidx = o; o = a256(o + 4 * (size_t)std::max(M.n_kp, 1LL));
    M.off_pred = o; o = a256(o + 16 * (size_t)std::max(qo, 1LL));
    M.off_qcell = o; o = a256(o + 4 * (size_t)std::max(qo, 1LL));
    M.off_best = o; o = a256(o + 4 * (size_t)std::max(qo, 1LL));
    M.off_b1 = o; o = a256(o + 8 * (size_t)std::max(qo, 1LL));
    M.off_kth = o; o = a256(o + 16 * (size_t)std::max(qo, 1LL));
    M.off_tkey = o; o = a256(o + 8 * (size_t)std::max(to, 1LL));
    M.off_tq = o; o = a256(o + 4 * (size_t)std::max(to, 1LL));
    M.off_cnt = o; o = a256(o + 16 * 4);
    M.off_outq = o; o = a256(o + 4 * (size_t)std::max(oo, 1LL));
    M.off_outt = o; o = a256(o + 4 * (size_t)std::max(oo, 1LL));
    M.off_outn = o; o = a256(o + 4 * (size_t)n_pair);
    M.total = o;
    return FSX_OK;
}

static MatchBatch batch_ptrs(const MatchPlanHost& M, char* ws, const uint8_t* desc, const double* xy, double radius,
                             int kc, double ratio) {
    MatchBatch b;
    memset(&b, 0, sizeof(b));
    b.desc = desc;
    b.xy = xy;
    b.img = (const MImage*)(ws + M.off_img);
    b.pair = (const MPair*)(ws + M.off_pair);
    b.n_img = (int)M.img.size();
    b.n_pair = (int)M.pair.size();
    b.n_query = M.n_query;
    b.n_kp = M.n_kp;
    b.n_cells = M.n_cells;
    b.n_targets = M.n_targets;
    b.radius = radius;
    b.r2 = radius * radius;
    b.ratio = ratio;
    b.kc = kc;
    b.norms = (unsigned*)(ws + M.off_norms);
    b.cell_start = (int*)(ws + M.off_cstart);
    b.cell_cur = (int*)(ws + M.off_ccur);
    b.cell_idx = (int*)(ws + M.off_cidx);
    b.pred = (double*)(ws + M.off_pred);
    b.qcell = (int*)(ws + M.off_qcell);
    b.best = (int*)(ws + M.off_best);
    b.b1 = (double*)(ws + M.off_b1);
    b.kth = (double*)(ws + M.off_kth);
    b.tkey = (unsigned long long*)(ws + M.off_tkey);
    b.tq = (int*)(ws + M.off_tq);
    b.counters = (int*)(ws + M.off_cnt);
    b.n_chunks = M.n_chunks;
    b.out_q = (int*)(ws + M.off_outq);
    b.out_t = (int*)(ws + M.off_outt);
    b.out_n = (int*)(ws + M.off_outn);
    return b;
}

static inline unsigned blocks_for(long long n) { return (unsigned)std::max(1LL, (n + 255) / 256); }

}  // namespace fsx

using namespace fsx;

extern "C" {

int fsx_match_workspace_bytes(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                              const int* pb, double radius, size_t* bytes, long long* out_capacity) {
    FSX_REQUIRE(bytes, "null pointer argument");
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, nullptr, nullptr, radius, M);
    if (rc) return rc;
    *bytes = M.total;
    if (out_capacity) *out_capacity = M.n_out;
    return FSX_OK;
}

int fsx_match_batch(int n_img, const int* n_kp, const int* W, const int* H, const uint8_t* desc_dev,
                    const double* xy_dev, int n_pair, const int* pa, const int* pb, const int* dense,
                    const double* Hm, double radius, int kc, double ratio, int form, void* ws, size_t ws_bytes,
                    long long* out_offsets, void* stage_events_v, void* stream) {
    FSX_REQUIRE(desc_dev && xy_dev && ws && pa && pb, "null pointer argument");
    FSX_REQUIRE(form == 0 || form == 1, "form must be 0 (plain FP32) or 1 (MFMA)");
    FSX_REQUIRE(kc >= 0 && kc <= 64, "kc must be in [0, 64]");
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, dense, Hm, radius, M);
    if (rc) return rc;
    FSX_REQUIRE(ws_bytes >= M.total, "match workspace too small: %zu < %zu", ws_bytes, M.total);
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)ws;
    MatchBatch b = batch_ptrs(M, w, desc_dev, xy_dev, radius, kc, ratio);
    rc = upload_async(w + M.off_img, M.img.data(), sizeof(MImage) * M.img.size(), st);
    if (rc) return rc;
    rc = upload_async(w + M.off_pair, M.pair.data(), sizeof(MPair) * M.pair.size(), st);
    if (rc) return rc;
    if (out_offsets)
        for (int p = 0; p < n_pair; p++) out_offsets[p] = M.pair[p].o_off;
    hipEvent_t* ev = (hipEvent_t*)stage_events_v;
    if (ev) FSX_HIP(hipEventRecord(ev[0], st));
    const long long nq = M.n_query;
    hipLaunchKernelGGL(k_m_reset, dim3(1024), dim3(256), 0, st, b);
    if (M.n_kp > 0) hipLaunchKernelGGL(k_m_norms_count, dim3(blocks_for(M.n_kp)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_m_grid_scan, dim3(n_img), dim3(256), 0, st, b);
    if (M.n_kp > 0) hipLaunchKernelGGL(k_m_grid_fill, dim3(blocks_for(M.n_kp)), dim3(256), 0, st, b);
    if (nq > 0) hipLaunchKernelGGL(k_m_predict, dim3(blocks_for(nq)), dim3(256), 0, st, b);
    if (nq > 0 && kc > 0) hipLaunchKernelGGL(k_m_knn, dim3(blocks_for(nq)), dim3(256), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[1], st));
    if (nq > 0) {
        if (form == 0) {
            hipLaunchKernelGGL(k_m_plain, dim3(blocks_for(nq)), dim3(256), 0, st, b);
        } else if (M.n_chunks > 0) {
            hipLaunchKernelGGL(k_m_mfma, dim3((unsigned)std::min<long long>(M.n_chunks, 1 << 20)), dim3(64), 0, st, b);
        }
        FSX_LAUNCH_CHECK();
    }
    if (ev) FSX_HIP(hipEventRecord(ev[2], st));
    if (nq > 0) hipLaunchKernelGGL(k_m_unique2, dim3(blocks_for(nq)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_m_compact, dim3(n_pair), dim3(256), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[3], st));
    return FSX_OK;
}

// Copies the results of the last fsx_match_batch on `ws` to host arrays (synchronises):
// counts[n_pair], then q/t of all pairs at the offsets fsx_match_batch reported.
int fsx_match_results(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                      const int* pb, double radius, const void* ws, int* counts, int* q, int* t, int* flags,
                      void* stream) {
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, nullptr, nullptr, radius, M);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const char* w = (const char*)ws;
    FSX_HIP(hipMemcpyAsync(counts, w + M.off_outn, 4 * (size_t)n_pair, hipMemcpyDeviceToHost, st));
    if (q) FSX_HIP(hipMemcpyAsync(q, w + M.off_outq, 4 * (size_t)M.n_out, hipMemcpyDeviceToHost, st));
    if (t) FSX_HIP(hipMemcpyAsync(t, w + M.off_outt, 4 * (size_t)M.n_out, hipMemcpyDeviceToHost, st));
    if (flags) FSX_HIP(hipMemcpyAsync(flags, w + M.off_cnt, 16, hipMemcpyDeviceToHost, st));
    FSX_HIP(hipStreamSynchronize(st));
    return FSX_OK;
}

// In-bounds check of the 32x32x16 BF16 MFMA operand/result lane layout this file assumes:
// A[i][k] = i + 32*k... computed for fixed small integer inputs; out: 64 lanes x 16 floats.
__global__ void k_mfma_layout_probe(float* out) {
    const int lane = threadIdx.x;
    bf16x8_t a, b;
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int k = (lane >> 5) * 8 + e;
        const float av = (float)((lane & 31) + (k & 3));        // A[row=lane%32][k]
        const float bv = (float)(((lane & 31) * 3 + k) & 7);     // B[k][col=lane%32]
        a[e] = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(av) >> 16));
        b[e] = __builtin_bit_cast(__bf16, (unsigned short)(__float_as_uint(bv) >> 16));
    }
    f32x16_t acc = {};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 16; r++) out[lane * 16 + r] = acc[r];
}

int fsx_mfma_layout_probe(float* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_mfma_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Per-query state of the last batch (synchronises): accepted best target (-1 = none), its
// distance b1, and the query's grid cell (-1 = predicted outside).  For tests and diagnostics.
int fsx_match_query_state(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                          const int* pb, double radius, const void* ws, int* best, double* b1, int* qcell,
                          void* stream) {
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, nullptr, nullptr, radius, M);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const char* w = (const char*)ws;
    if (M.n_query > 0) {
        if (best) FSX_HIP(hipMemcpyAsync(best, w + M.off_best, 4 * (size_t)M.n_query, hipMemcpyDeviceToHost, st));
        if (b1) FSX_HIP(hipMemcpyAsync(b1, w + M.off_b1, 8 * (size_t)M.n_query, hipMemcpyDeviceToHost, st));
        if (qcell) FSX_HIP(hipMemcpyAsync(qcell, w + M.off_qcell, 4 * (size_t)M.n_query, hipMemcpyDeviceToHost, st));
    }
    FSX_HIP(hipStreamSynchronize(st));
    return FSX_OK;
}

// Workspace layout (byte offsets) for tests and diagnostics: [0] norms [1] cell_start [2] cell_cur
// [3] cell_idx [4] pred [5] qcell [6] best [7] b1 [8] kth [9] tkey [10] tq [11] counters
// [12] out_q [13] out_t [14] out_n [15] n_chunks [16] total
int fsx_match_layout(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                     const int* pb, double radius, long long* offs) {
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, nullptr, nullptr, radius, M);
    if (rc) return rc;
    const size_t v[17] = {M.off_norms, M.off_cstart, M.off_ccur, M.off_cidx, M.off_pred, M.off_qcell, M.off_best,
                          M.off_b1, M.off_kth, M.off_tkey, M.off_tq, M.off_cnt, M.off_outq, M.off_outt, M.off_outn,
                          (size_t)M.n_chunks, M.total};
    for (int i = 0; i < 17; i++) offs[i] = (long long)v[i];
    return FSX_OK;
}

}  // extern "C"
