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
// uint8 RootSIFT descriptors make every d an exact integer < 2^24, so the float32 form and
// the integer MFMA form are exact.
//
// Three launches per batch:
//   setup:       one block per image buckets its keypoints into a grid of radius-sized cells
//                (tile-major numbering) and writes the cell-sorted index and positions, so a
//                cell's candidates are one contiguous range; the remaining blocks compute the
//                descriptor norms and reset the per-target keys;
//   contraction: plain: one thread per query, FP32 gather + dot products on VALU (the
//                baseline form); mfma: 32 cell-consecutive queries per wave, the union of
//                their candidate cells contracted on MFMA with the prior-window mask and the
//                top-2 selection fused into the epilogue.  The MFMA operands are the bytes
//                shifted by -128 (u8 ^ 0x80 = signed int8): sum_k (a_k - b_k)^2 is unchanged by
//                a common shift, so d = |a'|^2 + |b'|^2 - 2 a'.b' with a 32x32x32 INT8 MFMA and
//                exact int32 accumulation.
//                Both predict inline and publish accepted matches with one 64-bit atomicMin
//                of (b1 float bits, query) per target, which is the one-query-per-target rule;
//   compact:     per pair, the surviving targets in target order -> (q, t) lists.
#include "fsx_common.h"
#include <math.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <vector>

namespace fsx {

struct MImage {
    long long off;      // first keypoint in the concatenated arrays (caller's and cell-sorted)
    int n;              // keypoints
    int W, H;           // image size (inside test on the B side)
    int gw, gh;         // spatial grid (cells of csz = radius * (1 + 2^-20) px covering
                        // [-2 csz, W + 2 csz) x [-2 csz, H + 2 csz))
    int tw4;            // tiles of 4x4 cells per tile row; cells are numbered tile-major
    int ncell;          // cell slots (whole tiles)
    long long cell_off; // first cell of this image in the cell arrays
    long long chunk_off;// first slot of this image in the chunk table
    int chunk_cap;      // chunk slots: ceil(n / 32) + tile rows
    int pad;
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
    const uint8_t* desc;          // [Ntot][128] RootSIFT uint8 (caller's order)
    const double* xy;             // [Ntot][2] positions (undistorted px)
    const MImage* img;
    const MPair* pair;
    int n_img, n_pair;
    long long n_query, n_kp, n_cells, n_targets, n_chunks;
    double radius, r2, ratio;
    double csz;                   // grid cell size
    int kc;
    int defer;                    // 1: the contraction leaves publication to k_m_publish
    unsigned* norms;              // [Ntot] sum of squares of the descriptor
    unsigned* snorms;             // [Ntot] sum of squares of the descriptor - 128
    // cell-sorted index: positions [off, off + n) of each image in grid-cell order
    double* s_xy;                 // [Ntot][2] positions
    int* s_orig;                  // [Ntot] the keypoint's index in its image
    int* cell_start;              // [n_cells + n_img] per image ncell + 1 offsets
    int* cell_cur;                // [n_cells] cursors of grids too large for LDS
    int2* chunks;                 // [sum chunk_cap] (first sorted position, length) of each MFMA
                                  // chunk: up to 32 keypoints of one tile row
    int* n_chunks_img;            // [n_img] chunks of each image
    // per query, in the caller's order (diagnostics)
    int* best;                    // accepted best target, -1 = none
    double* b1;                   // its distance
    int* qcell;                   // cell of the prediction in b's grid, -1 = not inside
    double* kth;                  // [n_query][2] kc-th nearest (d2, idx bits) key, kc > 0 only
    int* cbest;                   // [n_query] best candidate before the ratio test, -1 = none
    float* cb12;                  // [n_query][2] its distance and the second smallest
    unsigned long long* tkey;     // [n_targets] min over accepted queries of (b1 bits << 32 | q)
    int* counters;                // [0..3] flags
    int* out_q;                   // [sum min(n_a, n_b)]
    int* out_t;
    int* out_n;                   // [n_pair]
};

// ------------------------------------------------------------------------------------------
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
    cx = (int)floor((x + 2.0 * B.csz) / B.csz);
    cy = (int)floor((y + 2.0 * B.csz) / B.csz);
    cx = cx < 0 ? 0 : (cx >= im.gw ? im.gw - 1 : cx);
    cy = cy < 0 ? 0 : (cy >= im.gh ? im.gh - 1 : cy);
}

// ------------------------------------------------------------------------------------------
// Setup

constexpr int SU_THREADS = 1024;
constexpr int SU_LDS_CELLS = 8192;      // grids up to this many cells count in LDS
constexpr int SU_KPT = 8;               // keypoints per thread held in registers

__global__ __launch_bounds__(SU_THREADS) void k_m_setup(const MatchBatch B) {
    __shared__ int s_cnt[SU_LDS_CELLS];
    __shared__ int s_w[SU_THREADS / 64];
    __shared__ int s_rowc[SU_THREADS];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if ((int)blockIdx.x >= B.n_img) {
        const long long base = (long long)(blockIdx.x - B.n_img) * SU_THREADS + tid;
        const long long stride = (long long)(gridDim.x - B.n_img) * SU_THREADS;
        for (long long t = base; t < B.n_targets; t += stride) B.tkey[t] = ~0ull;
        for (long long k = base; k < B.n_kp; k += stride) {
            const uint4* d = (const uint4*)(B.desc + k * 128);
            unsigned s = 0, s8 = 0;
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const uint4 v = d[i];
                const unsigned w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
                for (int h = 0; h < 4; h++) {
                    s = __builtin_amdgcn_udot4(w[h], w[h], s, false);
                    const int x = (int)(w[h] ^ 0x80808080u);
                    s8 = (unsigned)__builtin_amdgcn_sdot4(x, x, (int)s8, false);
                }
            }
            B.norms[k] = s;
            B.snorms[k] = s8;
        }
        if (base < 4) B.counters[base] = 0;
        return;
    }
    const int ii = blockIdx.x;
    const MImage im = B.img[ii];      // by value: loaded once, no reloads after stores
    const int nc = im.ncell;
    int* cnt = nc <= SU_LDS_CELLS ? s_cnt : B.cell_cur + im.cell_off;   // only this block uses it
    for (int c = tid; c < nc; c += SU_THREADS) cnt[c] = 0;
    // keypoints k = tid + SU_THREADS * r, r < SU_KPT, stay in registers between the passes
    // (all loads in flight at once); larger images take the loop below for the rest
    double kx[SU_KPT], ky[SU_KPT];
    int kcell[SU_KPT];
#pragma unroll
    for (int r = 0; r < SU_KPT; r++) {
        const int k = tid + SU_THREADS * r;
        kx[r] = k < im.n ? B.xy[2 * (im.off + k)] : 0.0;
        ky[r] = k < im.n ? B.xy[2 * (im.off + k) + 1] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SU_KPT; r++) {
        const int k = tid + SU_THREADS * r;
        int cx, cy;
        cell_of(B, im, kx[r], ky[r], cx, cy);
        kcell[r] = cell_id(im, cx, cy);
        if (k < im.n) atomicAdd(&cnt[kcell[r]], 1);
    }
    for (int k = tid + SU_THREADS * SU_KPT; k < im.n; k += SU_THREADS) {
        int cx, cy;
        cell_of(B, im, B.xy[2 * (im.off + k)], B.xy[2 * (im.off + k) + 1], cx, cy);
        atomicAdd(&cnt[cell_id(im, cx, cy)], 1);
    }
    __syncthreads();
    // exclusive scan of the counts -> cell_start; the counts become insertion cursors
    int* start = B.cell_start + im.cell_off + ii;
    int carry = 0;
    for (int base = 0; base < nc; base += SU_THREADS) {
        const int i = base + tid;
        const int v = i < nc ? cnt[i] : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w2 = 0; w2 < SU_THREADS / 64; w2++) {
            pre += w2 < wv ? s_w[w2] : 0;
            tot += s_w[w2];
        }
        if (i < nc) {
            start[i] = carry + pre + x - v;
            cnt[i] = carry + pre + x - v;
        }
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) start[nc] = carry;
    __syncthreads();
    // MFMA chunk table: each tile row (4 cell rows, tw4 tiles) is cut into runs of <= 32
    // keypoints, so a chunk never wraps from the right edge of the grid to the left edge.
    // cnt[] holds the cell starts here (before the scatter advances them).
    {
        const int th4 = nc / (16 * im.tw4);
        int total_chunks = 0;
        for (int r0 = 0; r0 < th4; r0 += SU_THREADS) {
            const int tr = r0 + tid;
            int nch = 0;
            if (tr < th4) {
                const int b0 = cnt[tr * 16 * im.tw4];
                const int b1 = tr + 1 < th4 ? cnt[(tr + 1) * 16 * im.tw4] : im.n;
                nch = (b1 - b0 + 31) >> 5;
            }
            int x = nch;
            for (int off = 1; off < 64; off <<= 1) {
                const int y = __shfl_up(x, off);
                if (lane >= off) x += y;
            }
            if (lane == 63) s_w[wv] = x;
            __syncthreads();
            int pre = 0, tot = 0;
            for (int w2 = 0; w2 < SU_THREADS / 64; w2++) {
                pre += w2 < wv ? s_w[w2] : 0;
                tot += s_w[w2];
            }
            // row tr's chunks start at chunk index first; each thread writes its row's chunks
            // only when the row is short, otherwise the rows are split below
            if (tr < th4) s_rowc[tr & (SU_THREADS - 1)] = total_chunks + pre + x - nch;
            __syncthreads();
            // all threads: chunk c of this row group -> its row by binary search over s_rowc
            const int nrow = min(SU_THREADS, th4 - r0);
            for (int c = tid; c < tot; c += SU_THREADS) {
                const int cg = total_chunks + c;
                int lo = 0, hi = nrow - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (s_rowc[mid] <= cg) lo = mid;
                    else hi = mid - 1;
                }
                const int trr = r0 + lo;
                const int b0 = cnt[trr * 16 * im.tw4];
                const int b1 = trr + 1 < th4 ? cnt[(trr + 1) * 16 * im.tw4] : im.n;
                const int beg = b0 + 32 * (cg - s_rowc[lo]);
                B.chunks[im.chunk_off + cg] = make_int2(beg, min(32, b1 - beg));
            }
            total_chunks += tot;
            __syncthreads();
        }
        if (tid == 0) B.n_chunks_img[ii] = total_chunks;
    }
    // scatter into the cell-sorted index (order inside a cell is arbitrary: every result is
    // defined by indices, not by visiting order)
#pragma unroll
    for (int r = 0; r < SU_KPT; r++) {
        const int k = tid + SU_THREADS * r;
        if (k < im.n) {
            const long long s = im.off + atomicAdd(&cnt[kcell[r]], 1);
            B.s_orig[s] = k;
            B.s_xy[2 * s] = kx[r];
            B.s_xy[2 * s + 1] = ky[r];
        }
    }
    for (int k = tid + SU_THREADS * SU_KPT; k < im.n; k += SU_THREADS) {
        const long long kg = im.off + k;
        const double x = B.xy[2 * kg], y = B.xy[2 * kg + 1];
        int cx, cy;
        cell_of(B, im, x, y, cx, cy);
        const long long s = im.off + atomicAdd(&cnt[cell_id(im, cx, cy)], 1);
        B.s_orig[s] = k;
        B.s_xy[2 * s] = x;
        B.s_xy[2 * s + 1] = y;
    }
}

// ------------------------------------------------------------------------------------------
// Shared pieces of the two contraction forms

// prediction (cv2.perspectiveTransform: w = (h6 x + h7 y) + h8, |w| > FLT_EPSILON ? v * (1/w)
// : 0) and the strict inside test of matching.py:119-120; returns the cell in b's grid or -1.
// Brute-force pairs need two keypoints on each side (matching.py:88-89).
__device__ __forceinline__ int predict_cell(const MatchBatch& B, const MPair& P, const MImage& ia, const MImage& ib,
                                            double x, double y, double& px, double& py) {
    px = 0.0;
    py = 0.0;
    if (P.dense) return (ib.n >= 2 && ia.n >= 2) ? 0 : -1;
    const double w = (x * P.Hm[6] + y * P.Hm[7]) + P.Hm[8];
    if (fabs(w) > 1.1920928955078125e-07) {
        const double iw = 1.0 / w;
        px = ((x * P.Hm[0] + y * P.Hm[1]) + P.Hm[2]) * iw;
        py = ((x * P.Hm[3] + y * P.Hm[4]) + P.Hm[5]) * iw;
    }
    if (px > -B.radius && px < ib.W + B.radius && py > -B.radius && py < ib.H + B.radius && ib.n > 0) {
        int cx, cy;
        cell_of(B, ib, px, py, cx, cy);
        return cell_id(ib, cx, cy);
    }
    return -1;
}

// spatial (d2, idx) key comparison for the kc-nearest rule
__device__ __forceinline__ bool sp_less(double d, int i, double d2, int i2) { return d < d2 || (d == d2 && i < i2); }

// kc-th nearest spatial key of a query (kc > 0): the candidate set is {in radius and
// (d2, idx) <= that key}; (r2, INT_MAX) when fewer than kc keypoints are in the radius.  The KS
// (>= kc) smallest keys are kept sorted in registers by a branch-free insertion network
// (static indices only: a list indexed at run time would live in scratch memory).
constexpr int KMAX = 64;

template <int KS>
__device__ void knn_threshold(const MatchBatch& B, const MPair& P, const MImage& ib, int cell, double px, double py,
                              double& td, int& ti) {
    td = B.r2;
    ti = 0x7fffffff;
    if (B.kc <= 0 || cell < 0 || P.dense) return;
    double ld[KS];
    int li[KS];
#pragma unroll
    for (int j = 0; j < KS; j++) {
        ld[j] = INFINITY;
        li[j] = 0x7fffffff;
    }
    int nl = 0;                                   // keypoints in the radius
    const int kc = B.kc < KS ? B.kc : KS;
    int cx0, cy0;
    cell_xy(ib, cell, cx0, cy0);
    const int* cs = B.cell_start + ib.cell_off + P.b;
    for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
        for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
            const int c = cell_id(ib, cx, cy);
            for (int s = cs[c]; s < cs[c + 1]; s++) {
                const long long sb = ib.off + s;
                const int t = B.s_orig[sb];
                const double dx = B.s_xy[2 * sb] - px, dy = B.s_xy[2 * sb + 1] - py;
                const double d = dx * dx + dy * dy;
                if (!(d < B.r2)) continue;
                nl++;
                if (!sp_less(d, t, ld[KS - 1], li[KS - 1])) continue;
#pragma unroll
                for (int j = KS - 1; j >= 1; j--) {
                    const bool lp = sp_less(d, t, ld[j - 1], li[j - 1]);
                    const bool lc = sp_less(d, t, ld[j], li[j]);
                    ld[j] = lp ? ld[j - 1] : (lc ? d : ld[j]);
                    li[j] = lp ? li[j - 1] : (lc ? t : li[j]);
                }
                const bool l0 = sp_less(d, t, ld[0], li[0]);
                ld[0] = l0 ? d : ld[0];
                li[0] = l0 ? t : li[0];
            }
        }
    if (nl >= kc) {
#pragma unroll
        for (int j = 0; j < KS; j++)
            if (j == kc - 1) {
                td = ld[j];
                ti = li[j];
            }
    }
}

// The same key by binned selection (the insertion network above costs O(KS) per inserted
// keypoint, and with kc ~ half the window's keypoints most of them are inserted).  Pass 1 counts
// the in-radius keypoints per distance bin, bin = min(floor(d * KNN_NB / r2), KNN_NB - 1):
// monotone in d, so each bin holds a contiguous run of the (d, idx) order; the bin holding the
// kc-th key and the count `below` it follow.  Pass 2 keeps the KNN_SMALL smallest (d, idx) keys
// of that bin in an exact register list; the key of rank kc - below in it is the kc-th overall.
// When that rank exceeds KNN_SMALL (clusters, equal distances) td = -1 marks the query for
// k_m_knn_exact (knn_threshold).  cnt: this thread's KNN_NB counters in LDS, KNN_STRIDE words
// apart.
constexpr int KNN_NB = 32;
constexpr int KNN_SMALL = 8;
constexpr int KNN_STRIDE = 256;

// The 3x3 cells around (cx0, cy0) as at most six runs of consecutive cell ids (a row's cells are
// consecutive inside a 4x4 tile of cell_id), so their keypoints are runs of the sorted order;
// f(s0, s1) is called for each run's sorted rows [s0, s1).
template <class F>
__device__ __forceinline__ void for_window_runs(const MImage& ib, const int* cs, int cx0, int cy0, F&& f) {
    const int xlo = max(cx0 - 1, 0), xhi = min(cx0 + 1, ib.gw - 1);
    const int xm = min(xhi, xlo | 3);             // last cell of xlo's tile within [xlo, xhi]
#pragma unroll
    for (int dy = -1; dy <= 1; dy++) {
        const int cy = cy0 + dy;
        if (cy < 0 || cy >= ib.gh) continue;
        f(cs[cell_id(ib, xlo, cy)], cs[cell_id(ib, xm, cy) + 1]);
        if (xm < xhi) f(cs[cell_id(ib, xm + 1, cy)], cs[cell_id(ib, xhi, cy) + 1]);
    }
}

// squared distances of sorted rows s .. s + 3 of run [s, s1) from (px, py); rows at or past s1
// read row s1 - 1 and give +inf
__device__ __forceinline__ void dist4(const MatchBatch& B, long long off, int s, int s1, double px, double py,
                                      double (&d)[4]) {
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const long long sb = off + min(s + k, s1 - 1);
        const double dx = B.s_xy[2 * sb] - px, dy = B.s_xy[2 * sb + 1] - py;
        d[k] = s + k < s1 ? dx * dx + dy * dy : INFINITY;
    }
}

__device__ void knn_threshold_binned(const MatchBatch& B, const MPair& P, const MImage& ib, int cell, double px,
                                     double py, double& td, int& ti, int* cnt) {
    td = B.r2;
    ti = 0x7fffffff;
    if (B.kc <= 0 || cell < 0 || P.dense) return;
#pragma unroll
    for (int b = 0; b < KNN_NB; b++) cnt[b * KNN_STRIDE] = 0;
    const double r2 = B.r2;
    const double inv = (double)KNN_NB / r2;
    const int kc = B.kc;
    int cx0, cy0;
    cell_xy(ib, cell, cx0, cy0);
    const int* cs = B.cell_start + ib.cell_off + P.b;
    const long long off = ib.off;
    int nl = 0;
    for_window_runs(ib, cs, cx0, cy0, [&](int s0, int s1) {
        for (int s = s0; s < s1; s += 4) {
            double d[4];
            dist4(B, off, s, s1, px, py, d);
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (d[k] < r2) {
                    nl++;
                    atomicAdd(&cnt[min((int)(d[k] * inv), KNN_NB - 1) * KNN_STRIDE], 1);
                }
        }
    });
    if (nl < kc) return;
    int below = 0, bs = KNN_NB - 1;
    for (int b = 0; b < KNN_NB; b++) {
        const int c = cnt[b * KNN_STRIDE];
        if (below + c >= kc) {
            bs = b;
            break;
        }
        below += c;
    }
    const int r = kc - below;                    // rank of the kc-th key inside bin bs, >= 1
    if (r > KNN_SMALL) {
        td = -1.0;
        return;
    }
    double ld[KNN_SMALL];
    int li[KNN_SMALL];
#pragma unroll
    for (int j = 0; j < KNN_SMALL; j++) {
        ld[j] = INFINITY;
        li[j] = 0x7fffffff;
    }
    for_window_runs(ib, cs, cx0, cy0, [&](int s0, int s1) {
        for (int s = s0; s < s1; s += 4) {
            double d[4];
            dist4(B, off, s, s1, px, py, d);
#pragma unroll
            for (int k = 0; k < 4; k++) {
                const double dk = d[k];
                if (!(dk < r2) || min((int)(dk * inv), KNN_NB - 1) != bs) continue;
                const int t = B.s_orig[off + s + k];
                if (!sp_less(dk, t, ld[KNN_SMALL - 1], li[KNN_SMALL - 1])) continue;
#pragma unroll
                for (int j = KNN_SMALL - 1; j >= 1; j--) {
                    const bool lp = sp_less(dk, t, ld[j - 1], li[j - 1]);
                    const bool lc = sp_less(dk, t, ld[j], li[j]);
                    ld[j] = lp ? ld[j - 1] : (lc ? dk : ld[j]);
                    li[j] = lp ? li[j - 1] : (lc ? t : li[j]);
                }
                const bool l0 = sp_less(dk, t, ld[0], li[0]);
                ld[0] = l0 ? dk : ld[0];
                li[0] = l0 ? t : li[0];
            }
        }
    });
#pragma unroll
    for (int j = 0; j < KNN_SMALL; j++)
        if (j == r - 1) {
            td = ld[j];
            ti = li[j];
        }
}

// kc > 0: the kc-th nearest key of every query, one thread per query (before the contraction),
// by binned selection; queries it marks (kth = -1) are decided by k_m_knn_exact.  Threads take
// the queries in the cell-sorted order of image a (k_m_setup's s_orig), so the predictions of a
// wave's queries lie close together in b and its lanes scan mostly the same candidate cells.
__global__ __launch_bounds__(256) void k_m_knn(const MatchBatch B) {
    const MPair P = B.pair[blockIdx.y];      // records by value: loaded once, never reloaded
    const MImage ia = B.img[P.a];
    const MImage ib = B.img[P.b];
    const int sq = blockIdx.x * 256 + threadIdx.x;
    if (sq >= ia.n) return;
    const int q = B.s_orig[ia.off + sq];
    const long long g = P.q_off + q;
    const long long ka = ia.off + q;
    double px, py, td;
    int ti;
    __shared__ int s_cnt[KNN_NB * KNN_STRIDE];
    const int cell = predict_cell(B, P, ia, ib, B.xy[2 * ka], B.xy[2 * ka + 1], px, py);
    knn_threshold_binned(B, P, ib, cell, px, py, td, ti, s_cnt + threadIdx.x);
    B.kth[2 * g] = td;
    B.kth[2 * g + 1] = __longlong_as_double((long long)ti);
}

// the queries k_m_knn marked: the insertion network over all their candidates; KS the list
// length, the smallest of 8, 16, 24, 32, 48, 64 that holds kc
template <int KS>
__global__ __launch_bounds__(256) void k_m_knn_exact(const MatchBatch B) {
    const MPair P = B.pair[blockIdx.y];
    const MImage ia = B.img[P.a];
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= ia.n) return;
    const long long g = P.q_off + q;
    if (!(B.kth[2 * g] < 0.0)) return;
    const MImage ib = B.img[P.b];
    const long long ka = ia.off + q;
    double px, py, td;
    int ti;
    const int cell = predict_cell(B, P, ia, ib, B.xy[2 * ka], B.xy[2 * ka + 1], px, py);
    knn_threshold<KS>(B, P, ib, cell, px, py, td, ti);
    B.kth[2 * g] = td;
    B.kth[2 * g + 1] = __longlong_as_double((long long)ti);
}

__device__ __forceinline__ void kth_of(const MatchBatch& B, long long g, double& td, int& ti) {
    td = B.r2;
    ti = 0x7fffffff;
    if (B.kc > 0) {
        td = B.kth[2 * g];
        ti = (int)__double_as_longlong(B.kth[2 * g + 1]);
    }
}

__device__ __forceinline__ bool in_window(const MatchBatch& B, double d, int t, double td, int ti) {
    if (!(d < B.r2)) return false;
    if (B.kc <= 0) return true;
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

// ratio test in float64 (matching.py:145-147) and the one-query-per-target key: the smallest
// (b1, q) wins, which is _unique's lexsort((dd, t)) first element per target
__device__ __forceinline__ void accept_query(const MatchBatch& B, long long g, const MPair& P, int q, float v1, int i1,
                                             float v2) {
    const bool ok = i1 >= 0 && sqrt((double)v1) < B.ratio * sqrt((double)v2);
    B.cbest[g] = i1;
    B.cb12[2 * g] = v1;
    B.cb12[2 * g + 1] = v2;
    B.best[g] = ok ? i1 : -1;
    B.b1[g] = ok ? (double)v1 : 0.0;
    if (ok && !B.defer) {
        const unsigned long long key = ((unsigned long long)(__float_as_uint(v1) & 0x7fffffffu) << 32) | (unsigned)q;
        atomicMin(&B.tkey[P.t_off + i1], key);
    }
}

// ------------------------------------------------------------------------------------------
// The plain FP32 GPU form: one thread per query; the query descriptor stays in registers as
// float32, each in-window candidate descriptor is gathered from global memory (16-byte loads)
// and the dot product runs on VALU in FP32: d = na + nb - 2 a.b (exact for uint8 inputs)

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

__global__ __launch_bounds__(256) void k_m_plain(const MatchBatch B) {
    const MPair P = B.pair[blockIdx.y];      // records by value: loaded once, never reloaded
    const MImage ia = B.img[P.a];
    const MImage ib = B.img[P.b];
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= ia.n) return;
    const long long g = P.q_off + q;
    const long long ka = ia.off + q;
    double px, py;
    const int cell = predict_cell(B, P, ia, ib, B.xy[2 * ka], B.xy[2 * ka + 1], px, py);
    B.qcell[g] = cell;
    if (cell < 0) {
        B.best[g] = -1;
        B.cbest[g] = -1;
        return;
    }
    double td;
    int ti;
    kth_of(B, g, td, ti);
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
        const long long kb = ib.off + t;
        const float dot = u8dot_f32(qa, (const uint4*)(B.desc + kb * 128));
        const float d = fmaxf((na + (float)B.norms[kb]) - 2.0f * dot, 0.0f);
        top2_push(d, t, v1, i1, v2);
    };
    if (P.dense) {
        for (int t = 0; t < ib.n; t++) eval(t);
    } else {
        int cx0, cy0;
        cell_xy(ib, cell, cx0, cy0);
        const int* cs = B.cell_start + ib.cell_off + P.b;
        for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
            for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
                const int c = cell_id(ib, cx, cy);
                for (int s = cs[c]; s < cs[c + 1]; s++) {
                    const long long sb = ib.off + s;
                    const double dx = B.s_xy[2 * sb] - px, dy = B.s_xy[2 * sb + 1] - py;
                    const int t = B.s_orig[sb];
                    if (in_window(B, dx * dx + dy * dy, t, td, ti)) eval(t);
                }
            }
    }
    accept_query(B, g, P, q, v1, i1, v2);
}

// ------------------------------------------------------------------------------------------
// MFMA form.  A chunk = 32 consecutive keypoints of the query image in its cell-sorted order,
// so the chunk's predictions land in a compact region of B (the prior is a smooth mapping).
// One wave per chunk: the union of the queries' 3x3 candidate cells is listed once, its
// candidates are contracted against the 32 queries on MFMA, 32 candidates per 32x32x32 INT8
// tile (4 instructions cover the 128 bytes), and the epilogue applies the prior-window mask
// and the top-2 update.  Lane l and lane l+32 hold query l; lane half h = l / 32 supplies
// bytes [32 s + 16 h, 32 s + 16 h + 16) of its row to instruction s (the same byte positions
// for both operands, so the instruction's internal k order does not matter).

typedef int i32x4_t __attribute__((ext_vector_type(4)));
typedef int i32x16_t __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));

constexpr int MF_CAP = 256;     // candidates per pass staged in LDS
constexpr int MF_MAXC = 256;    // listed candidate cells per chunk

// 16 descriptor bytes as signed int8 (byte - 128)
__device__ __forceinline__ i32x4_t s8x16(uint4 v) {
    i32x4_t r;
    r[0] = (int)(v.x ^ 0x80808080u);
    r[1] = (int)(v.y ^ 0x80808080u);
    r[2] = (int)(v.z ^ 0x80808080u);
    r[3] = (int)(v.w ^ 0x80808080u);
    return r;
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
__device__ void query_direct(const MatchBatch& B, const MPair& P, const MImage& ib, long long ka, int cell,
                             double qpx, double qpy, double td, int ti, float& v1, int& i1, float& v2) {
    v1 = INFINITY;
    v2 = INFINITY;
    i1 = -1;
    const float qn = (float)B.norms[ka];
    const uint4* a4 = (const uint4*)(B.desc + ka * 128);
    int cx0, cy0;
    cell_xy(ib, cell, cx0, cy0);
    const int* cs = B.cell_start + ib.cell_off + P.b;
    for (int cy = max(cy0 - 1, 0); cy <= min(cy0 + 1, ib.gh - 1); cy++)
        for (int cx = max(cx0 - 1, 0); cx <= min(cx0 + 1, ib.gw - 1); cx++) {
            const int cc = cell_id(ib, cx, cy);
            for (int s = cs[cc]; s < cs[cc + 1]; s++) {
                const long long sb = ib.off + s;
                const int t = B.s_orig[sb];
                const double dx = B.s_xy[2 * sb] - qpx, dy = B.s_xy[2 * sb + 1] - qpy;
                if (!in_window(B, dx * dx + dy * dy, t, td, ti)) continue;
                const uint4* b4 = (const uint4*)(B.desc + (ib.off + t) * 128);
                unsigned dot = 0;
#pragma unroll
                for (int k = 0; k < 8; k++) {
                    const uint4 x = a4[k], y = b4[k];
                    dot = __builtin_amdgcn_udot4(x.x, y.x, dot, false);
                    dot = __builtin_amdgcn_udot4(x.y, y.y, dot, false);
                    dot = __builtin_amdgcn_udot4(x.z, y.z, dot, false);
                    dot = __builtin_amdgcn_udot4(x.w, y.w, dot, false);
                }
                const float d = fmaxf((qn + (float)B.norms[ib.off + t]) - 2.0f * (float)dot, 0.0f);
                top2_push(d, t, v1, i1, v2);
            }
        }
}

// One 64-lane workgroup per MFMA chunk.  PERSIST = false: the grid is (chunk, pair);
// PERSIST = true: a fixed number of persistent workgroups stride over the (pair, chunk) items,
// pair-major, max_chunks items per pair (chunks past an image's count are skipped).
template <bool PERSIST>
__global__ __launch_bounds__(64) void k_m_mfma(const MatchBatch B, int max_chunks) {
    __shared__ int s_t[MF_CAP];
    __shared__ double s_cx[MF_CAP], s_cy[MF_CAP];
    __shared__ __attribute__((aligned(16))) int s_cn[MF_CAP];
    __shared__ unsigned s_mark[32];
    __shared__ int s_cells[MF_MAXC], s_coff[MF_MAXC + 1], s_cst[MF_MAXC];
    __shared__ float s_v1[64], s_v2[64];
    __shared__ int s_i1[64];
    const int lane = threadIdx.x;
    const int qc = lane & 31;
    const int n_items = PERSIST ? max_chunks * B.n_pair : 1;       // < 2^31 (checked at launch)
    for (int item = PERSIST ? (int)blockIdx.x : 0; item < n_items; item += PERSIST ? (int)gridDim.x : 1) {
        const int chunk = PERSIST ? item % max_chunks : (int)blockIdx.x;
        const MPair P = B.pair[PERSIST ? item / max_chunks : (int)blockIdx.y];   // records by value
        const MImage ia = B.img[P.a];
        const MImage ib = B.img[P.b];
        if (chunk >= B.n_chunks_img[P.a]) continue;      // block-uniform
        const int2 chk = B.chunks[ia.chunk_off + chunk];
        const int kk = chk.x + qc;
        const bool valid = qc < chk.y;
        const long long sa = ia.off + (valid ? kk : 0);
        const int q = valid ? B.s_orig[sa] : -1;
        const long long g = P.q_off + (q >= 0 ? q : 0);
        double qpx = 0.0, qpy = 0.0;
        const int cell = valid ? predict_cell(B, P, ia, ib, B.s_xy[2 * sa], B.s_xy[2 * sa + 1], qpx, qpy) : -1;
        const bool act = cell >= 0;
        if (lane < 32) {
            s_mark[lane] = 0u;
            if (valid) B.qcell[g] = cell;
        }
        __syncthreads();
        if (__ballot(act) == 0ull) {       // wave-uniform: nothing to match in this chunk
            if (lane < 32 && valid) {
                B.best[g] = -1;
                B.cbest[g] = -1;
            }
            __syncthreads();               // s_mark is cleared again by the next item
            continue;
        }
        double td = B.r2;
        int ti = 0x7fffffff;
        if (act) kth_of(B, g, td, ti);
        // window origin: min cell coordinates of the active queries, minus one
        int cxq = 1 << 30, cyq = 1 << 30;
        if (act) cell_xy(ib, cell, cxq, cyq);
        int cx_lo = cxq, cy_lo = cyq;
        for (int off = 32; off >= 1; off >>= 1) {
            cx_lo = min(cx_lo, __shfl_xor(cx_lo, off));
            cy_lo = min(cy_lo, __shfl_xor(cy_lo, off));
        }
        cx_lo = max(cx_lo - 1, 0);
        cy_lo = max(cy_lo - 1, 0);
        bool overflow = false;
        if (!P.dense && lane < 32 && act) {
            for (int cy = max(cyq - 1, 0); cy <= min(cyq + 1, ib.gh - 1); cy++)
                for (int cx = max(cxq - 1, 0); cx <= min(cxq + 1, ib.gw - 1); cx++) {
                    const int bx = cx - cx_lo, by = cy - cy_lo;
                    if (bx < 32 && by < 32) atomicOr(&s_mark[by], 1u << bx);
                    else overflow = true;
                }
        }
        bool any_over = __ballot(overflow) != 0ull;
        __syncthreads();
        // list the marked cells (window-row order) and their candidate offsets
        int total = 0;
        int ncell = 0;
        const int* cs = B.cell_start + ib.cell_off + P.b;
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
                int carry = 0;
                for (int c0 = 0; c0 < ncell; c0 += 64) {
                    const int c = c0 + lane;
                    const int c_lo = c < ncell ? cs[s_cells[c]] : 0;
                    const int n = c < ncell ? cs[s_cells[c] + 1] - c_lo : 0;
                    int tot;
                    const int ex = wave_excl_scan(n, lane, tot);
                    if (c < ncell) {
                        s_coff[c] = carry + ex;
                        s_cst[c] = c_lo;
                    }
                    carry += tot;
                }
                if (lane == 0) s_coff[ncell] = carry;
                total = carry;
                __syncthreads();
            }
        }
        // this lane's query column over the candidate passes
        float v1 = INFINITY, v2 = INFINITY;
        int i1 = -1;
        const long long ka = ia.off + (q >= 0 ? q : 0);
        i32x4_t qop[4];
        const int qn = act ? (int)B.snorms[ka] : 0;
        {
            const uint4* qa = (const uint4*)(B.desc + ka * 128);
#pragma unroll
            for (int s = 0; s < 4; s++) qop[s] = s8x16(act ? qa[2 * s + (lane >> 5)] : make_uint4(0, 0, 0, 0));
        }
        if (!any_over) {
            for (int base = 0; base < total; base += MF_CAP) {
                const int np = min(MF_CAP, total - base);
                // stage the pass's candidates: original index and position (up to MF_CAP / 64
                // per lane, all loads issued before any LDS store)
                {
                    long long sbv[MF_CAP / 64];
#pragma unroll
                    for (int u = 0; u < MF_CAP / 64; u++) {
                        const int j = lane + 64 * u;
                        sbv[u] = -1;
                        if (j < np && !P.dense) {
                            // listed cell holding candidate position base + j: the last cell
                            // whose offset is <= pos (binary search over s_coff[0..ncell))
                            const int pos = base + j;
                            int lo = 0, hi = ncell - 1;
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (s_coff[mid] <= pos) lo = mid;
                                else hi = mid - 1;
                            }
                            sbv[u] = ib.off + s_cst[lo] + (pos - s_coff[lo]);
                        }
                    }
                    double cxv[MF_CAP / 64], cyv[MF_CAP / 64];
                    int tv[MF_CAP / 64];
#pragma unroll
                    for (int u = 0; u < MF_CAP / 64; u++) {
                        const long long sb = sbv[u];
                        cxv[u] = sb >= 0 ? B.s_xy[2 * sb] : 0.0;
                        cyv[u] = sb >= 0 ? B.s_xy[2 * sb + 1] : 0.0;
                        tv[u] = sb >= 0 ? B.s_orig[sb] : base + lane + 64 * u;    // dense: identity
                    }
#pragma unroll
                    for (int u = 0; u < MF_CAP / 64; u++) {
                        const int j = lane + 64 * u;
                        if (j < np) {
                            s_t[j] = tv[u];
                            s_cx[j] = cxv[u];
                            s_cy[j] = cyv[u];
                        }
                    }
                }
                __syncthreads();
                // tiles of 32 candidates; the next tile's bytes are loaded while this one is
                // contracted and reduced
                uint4 cur[4], nxt[4];
                auto load_tile = [&](int t0, uint4* dst) {
                    const int tr = t0 + (lane & 31);
                    const int tc = tr < np ? s_t[tr] : -1;
                    const uint4* cb = (const uint4*)(B.desc + (ib.off + (tc >= 0 ? tc : 0)) * 128);
#pragma unroll
                    for (int s = 0; s < 4; s++) dst[s] = tc >= 0 ? cb[2 * s + (lane >> 5)] : make_uint4(0, 0, 0, 0);
                };
                load_tile(0, cur);
                for (int t0 = 0; t0 < np; t0 += 32) {
                    if (t0 + 32 < np) load_tile(t0 + 32, nxt);
                    i32x16_t acc = {};
                    int nrm = 0;     // |b'|^2 of this lane's half row, completed with lane ^ 32
#pragma unroll
                    for (int s = 0; s < 4; s++) {
                        const i32x4_t a = s8x16(cur[s]);
                        acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, qop[s], acc, 0, 0, 0);
#pragma unroll
                        for (int w = 0; w < 4; w++) nrm = __builtin_amdgcn_sdot4(a[w], a[w], nrm, false);
                    }
                    nrm += __shfl_xor(nrm, 32);
                    if (lane < 32) s_cn[t0 + lane] = nrm;
                    __syncthreads();
                    if (act) {
#pragma unroll
                        for (int i = 0; i < 4; i++) {
                            const int4 cn4 = *(const int4*)&s_cn[t0 + i * 8 + (lane >> 5) * 4];
                            const int cnv[4] = {cn4.x, cn4.y, cn4.z, cn4.w};
#pragma unroll
                            for (int j = 0; j < 4; j++) {
                                const int r = t0 + i * 8 + (lane >> 5) * 4 + j;
                                if (r >= np) continue;
                                // exact: |a'|^2 + |b'|^2 - 2 a'.b' = sum (a - b)^2 < 2^24
                                const float d = (float)(qn + cnv[j] - 2 * acc[i * 4 + j]);
                                // a distance above the current second best cannot change the
                                // top-2, whatever its window membership
                                if (d > v2) continue;
                                const int t = s_t[r];
                                if (!P.dense) {
                                    const double dx = s_cx[r] - qpx, dy = s_cy[r] - qpy;
                                    if (!in_window(B, dx * dx + dy * dy, t, td, ti)) continue;
                                }
                                top2_push(d, t, v1, i1, v2);
                            }
                        }
                    }
#pragma unroll
                    for (int s = 0; s < 4; s++) cur[s] = nxt[s];
                }
                __syncthreads();
            }
        }
        // merge the two half-waves of each query column
        s_v1[lane] = v1;
        s_v2[lane] = v2;
        s_i1[lane] = i1;
        __syncthreads();
        if (lane < 32 && valid) {
            if (!act) {
                B.best[g] = -1;
                B.cbest[g] = -1;
            } else {
                if (any_over) {
                    query_direct(B, P, ib, ka, cell, qpx, qpy, td, ti, v1, i1, v2);
                } else {
                    const float w1 = s_v1[lane + 32], w2 = s_v2[lane + 32];
                    const int j1 = s_i1[lane + 32];
                    if (j1 >= 0) top2_push(w1, j1, v1, i1, v2);
                    v2 = fminf(v2, w2);
                }
                accept_query(B, g, P, q, v1, i1, v2);
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------------
// Learned-filter inputs: one row of 288 bf16 per query (its best candidate before the ratio
// test), 64 threads per row: lanes 0..31 write the descriptor terms (4 bins each), lane 32 the
// geometric terms, the rest the zero padding.

__device__ __forceinline__ unsigned short f_bf16(float x) {
    const unsigned b = __float_as_uint(x);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

// one 288-wide row per query, written by the 64 lanes of a wave
__device__ void write_filter_row(const MatchBatch& B, const MPair& P, const MImage& ia, const MImage& ib, int q,
                                 long long g, const float* __restrict__ size, const float* __restrict__ angle,
                                 unsigned short* __restrict__ row, int lane) {
    const int t = B.cbest[g];
    if (t < 0) {
        for (int k = lane; k < 288; k += 64) row[k] = 0;
        return;
    }
    const long long ka = ia.off + q, kb = ib.off + t;
    if (lane < 32) {
        const unsigned a4 = ((const unsigned*)(B.desc + ka * 128))[lane];
        const unsigned b4 = ((const unsigned*)(B.desc + kb * 128))[lane];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const float av = (float)((a4 >> (8 * u)) & 0xffu), bv = (float)((b4 >> (8 * u)) & 0xffu);
            row[4 * lane + u] = f_bf16(fabsf(av - bv) * (1.0f / 255.0f));
            row[128 + 4 * lane + u] = f_bf16(av * bv * (1.0f / 65025.0f));
        }
    } else if (lane == 32) {
        const double x = B.xy[2 * ka], y = B.xy[2 * ka + 1];
        double px, py;
        predict_cell(B, P, ia, ib, x, y, px, py);
        const float b1 = B.cb12[2 * g], b2 = B.cb12[2 * g + 1];
        const float r = (float)B.radius;
        float dang = angle[kb] - angle[ka];
        dang = dang >= 180.0f ? dang - 360.0f : (dang < -180.0f ? dang + 360.0f : dang);
        const float sa = size[ka], sb = size[kb];
        const float v[8] = {(float)((px - B.xy[2 * kb]) / B.radius), (float)((py - B.xy[2 * kb + 1]) / B.radius),
                            sqrtf(b1) / 1024.0f, sqrtf(fminf(b2, 4e6f)) / 1024.0f,
                            isinf(b2) ? 0.0f : sqrtf(b1 / fmaxf(b2, 1e-30f)),
                            (sa > 0.0f && sb > 0.0f) ? logf(sb / sa) : 0.0f, dang / 180.0f, r / 100.0f};
#pragma unroll
        for (int u = 0; u < 8; u++) row[256 + u] = f_bf16(v[u]);
    } else if (lane < 57) {
        row[264 + (lane - 33)] = 0;
    }
}

__global__ __launch_bounds__(256) void k_m_features(const MatchBatch B, const float* __restrict__ size,
                                                    const float* __restrict__ angle, unsigned short* __restrict__ X) {
    const MPair P = B.pair[blockIdx.y];
    const MImage ia = B.img[P.a];
    const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (q >= ia.n) return;
    const long long g = P.q_off + q;
    write_filter_row(B, P, ia, B.img[P.b], q, g, size, angle, X + g * 288, threadIdx.x & 63);
}

// ------------------------------------------------------------------------------------------
// Learned-filter stage (contraction with defer = 1, then these, then compaction):
//   k_m_accepted: the queries that passed the ratio test, as (global query, pair) entries of a
//                 list in arbitrary order (one atomic slot each);
//   k_m_features_list: the filter's input row of each list entry (the row k_m_features writes);
//   k_m_publish: list entries whose score is >= tau publish their (b1, query) key into the
//                per-target minimum (the one-query-per-target rule among the kept queries);
//                the others are marked rejected (best = -1).  score = null keeps every entry.

__global__ __launch_bounds__(256) void k_m_accepted(const MatchBatch B, int2* __restrict__ list,
                                                    int* __restrict__ count) {
    const MPair P = B.pair[blockIdx.y];
    const int q = blockIdx.x * 256 + threadIdx.x;
    if (q >= B.img[P.a].n) return;
    const long long g = P.q_off + q;
    if (B.best[g] >= 0) {
        const int slot = atomicAdd(count, 1);
        list[slot] = make_int2((int)g, (int)blockIdx.y);
    }
}

__global__ __launch_bounds__(256) void k_m_features_list(const MatchBatch B, const float* __restrict__ size,
                                                         const float* __restrict__ angle, const int2* __restrict__ list,
                                                         const int* __restrict__ count, unsigned short* __restrict__ X) {
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= *count) return;
    const int2 e = list[r];
    const MPair P = B.pair[e.y];
    write_filter_row(B, P, B.img[P.a], B.img[P.b], (int)(e.x - P.q_off), e.x, size, angle, X + (long long)r * 288,
                     threadIdx.x & 63);
}

__global__ __launch_bounds__(256) void k_m_publish(const MatchBatch B, const int2* __restrict__ list,
                                                   const int* __restrict__ count, const float* __restrict__ score,
                                                   float tau) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= *count) return;
    const int2 e = list[r];
    const MPair P = B.pair[e.y];
    const long long g = e.x;
    const int t = B.best[g];
    if (score && !(score[r] >= tau)) {
        B.best[g] = -1;
        B.b1[g] = 0.0;
        return;
    }
    const float v1 = (float)B.b1[g];
    const unsigned long long key =
        ((unsigned long long)(__float_as_uint(v1) & 0x7fffffffu) << 32) | (unsigned)(g - P.q_off);
    atomicMin(&B.tkey[P.t_off + t], key);
}

// ------------------------------------------------------------------------------------------
// Compact: one block per pair, winners in target order -> (q, t) lists

constexpr int FSX_MATCH_INFO_LEN = 64;   // launch_info entries (header + the batch record)
constexpr int CP_THREADS = 1024;
constexpr int CP_TPT = 8;          // targets per thread

__global__ __launch_bounds__(CP_THREADS) void k_m_compact(const MatchBatch B) {
    __shared__ int s_w[CP_THREADS / 64];
    const MPair P = B.pair[blockIdx.x];
    const int nb = B.img[P.b].n;
    const unsigned long long* tk = B.tkey + P.t_off;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int carry = 0;
    for (int base0 = 0; base0 < nb; base0 += CP_THREADS * CP_TPT) {
        // thread tid owns targets base0 + tid * CP_TPT + [0, CP_TPT): all keys loaded at once
        unsigned long long key[CP_TPT];
        int cnt = 0;
#pragma unroll
        for (int r = 0; r < CP_TPT; r++) {
            const int t = base0 + tid * CP_TPT + r;
            key[r] = t < nb ? tk[t] : ~0ull;
            cnt += key[r] != ~0ull ? 1 : 0;
        }
        int x = cnt;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w2 = 0; w2 < CP_THREADS / 64; w2++) {
            pre += w2 < wv ? s_w[w2] : 0;
            tot += s_w[w2];
        }
        long long o = P.o_off + carry + pre + x - cnt;
#pragma unroll
        for (int r = 0; r < CP_TPT; r++) {
            if (key[r] != ~0ull) {
                B.out_q[o] = (int)(unsigned)(key[r] & 0xffffffffull);
                B.out_t[o] = base0 + tid * CP_TPT + r;
                o++;
            }
        }
        carry += tot;
        __syncthreads();
    }
    if (tid == 0) B.out_n[blockIdx.x] = carry;
}

// The last run's per-pair lists packed on the device: pair p's list (out_n[p] entries at its
// output slot o_off) to rows poff[p] .. of q_out / t_out; a list that would pass row `cap` is
// not written.
__global__ __launch_bounds__(256) void k_m_pack(const MatchBatch B, const long long* __restrict__ poff, long long cap,
                                               int* __restrict__ q_out, int* __restrict__ t_out) {
    const int p = blockIdx.x;
    const long long o = B.pair[p].o_off;
    const int n = B.out_n[p];
    const long long d = poff[p];
    if (d < 0 || n < 0 || d + n > cap) return;
    for (int i = threadIdx.x; i < n; i += 256) {
        q_out[d + i] = B.out_q[o + i];
        t_out[d + i] = B.out_t[o + i];
    }
}

// ------------------------------------------------------------------------------------------
// Host side

static inline size_t a256(size_t x) { return (x + 255) & ~(size_t)255; }

// Cells slightly larger than the radius: two points closer than the radius then always land in
// the same or adjacent cells despite the rounding of floor((x + 2 csz) / csz) (a cell
// difference of 2 needs a computed quotient difference above 1, i.e. a true distance above
// csz * (1 - 4 ulp) > radius), so the 3x3 cells around a prediction hold its whole window.
static inline double cell_size(double radius) { return radius * (1.0 + 0x1p-20); }

// A host table held in a std::vector, or in fixed storage the caller provides (the
// no-allocation entry point plans into arrays on its stack).
template <class T>
struct HostTab {
    std::vector<T> store;
    T* ext = nullptr;
    size_t cap = 0, n = 0;
    bool resize(size_t k) {
        if (ext) {
            if (k > cap) return false;
        } else {
            store.resize(k);
        }
        n = k;
        return true;
    }
    T& operator[](size_t i) { return ext ? ext[i] : store[i]; }
    const T& operator[](size_t i) const { return ext ? ext[i] : store[i]; }
    size_t size() const { return n; }
    T* data() { return ext ? ext : store.data(); }
};

struct MatchPlanHost {
    HostTab<MImage> img;
    HostTab<MPair> pair;
    long long n_kp = 0, n_query = 0, n_cells = 0, n_targets = 0, n_out = 0, n_chunks = 0, n_chunk_slots = 0;
    size_t off_chunks, off_nchunks, off_batch, off_img, off_pair, off_norms, off_snorms, off_sxy, off_sorig, off_cstart, off_ccur, off_best, off_b1,
        off_qcell, off_kth, off_cbest, off_cb12, off_tkey, off_cnt, off_outq, off_outt, off_outn, total;
};

static int match_plan(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                      const int* pb, const int* dense, const double* Hm, double radius, MatchPlanHost& M) {
    FSX_REQUIRE(n_img >= 1 && n_pair >= 1, "need images and pairs");
    FSX_REQUIRE(radius > 0, "radius must be > 0");
    FSX_REQUIRE(M.img.resize(n_img), "too many images (%d) for this entry point", n_img);
    long long off = 0, coff = 0, choff = 0;
    for (int i = 0; i < n_img; i++) {
        FSX_REQUIRE(n_kp[i] >= 0 && W[i] > 0 && H[i] > 0, "bad image %d", i);
        MImage& m = M.img[i];
        m.off = off;
        m.n = n_kp[i];
        m.W = W[i];
        m.H = H[i];
        const double csz = cell_size(radius);
        const double gw = ceil((W[i] + 4.0 * csz) / csz), gh = ceil((H[i] + 4.0 * csz) / csz);
        FSX_REQUIRE(gw * gh < 1e8, "image %d: radius %g gives a grid of %.0f cells", i, radius, gw * gh);
        m.gw = (int)gw;
        m.gh = (int)gh;
        m.tw4 = (m.gw + 3) / 4;
        m.ncell = m.tw4 * ((m.gh + 3) / 4) * 16;
        m.cell_off = coff;
        m.chunk_off = choff;
        m.chunk_cap = (n_kp[i] + 31) / 32 + (m.gh + 3) / 4;
        off += n_kp[i];
        coff += m.ncell;
        choff += m.chunk_cap;
    }
    M.n_kp = off;
    M.n_cells = coff;
    M.n_chunk_slots = choff;
    FSX_REQUIRE(M.pair.resize(n_pair), "too many pairs (%d) for this entry point", n_pair);
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
    const size_t nk = (size_t)std::max(M.n_kp, 1LL), nq = (size_t)std::max(qo, 1LL);
    size_t o = 0;
    M.off_batch = o; o = a256(o + sizeof(MatchBatch));
    M.off_img = o; o = a256(o + sizeof(MImage) * n_img);
    M.off_pair = o; o = a256(o + sizeof(MPair) * n_pair);
    M.off_norms = o; o = a256(o + 4 * nk);
    M.off_snorms = o; o = a256(o + 4 * nk);
    M.off_sxy = o; o = a256(o + 16 * nk);
    M.off_sorig = o; o = a256(o + 4 * nk);
    M.off_cstart = o; o = a256(o + 4 * (size_t)(M.n_cells + n_img));
    M.off_ccur = o; o = a256(o + 4 * (size_t)M.n_cells);
    M.off_chunks = o; o = a256(o + 8 * (size_t)std::max(M.n_chunk_slots, 1LL));
    M.off_nchunks = o; o = a256(o + 4 * (size_t)n_img);
    M.off_best = o; o = a256(o + 4 * nq);
    M.off_b1 = o; o = a256(o + 8 * nq);
    M.off_qcell = o; o = a256(o + 4 * nq);
    M.off_kth = o; o = a256(o + 16 * nq);
    M.off_cbest = o; o = a256(o + 4 * nq);
    M.off_cb12 = o; o = a256(o + 8 * nq);
    M.off_tkey = o; o = a256(o + 8 * (size_t)std::max(to, 1LL));
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
    b.n_chunks = M.n_chunks;
    b.radius = radius;
    b.r2 = radius * radius;
    b.csz = cell_size(radius);
    b.ratio = ratio;
    b.kc = kc;
    b.norms = (unsigned*)(ws + M.off_norms);
    b.snorms = (unsigned*)(ws + M.off_snorms);
    b.s_xy = (double*)(ws + M.off_sxy);
    b.s_orig = (int*)(ws + M.off_sorig);
    b.cell_start = (int*)(ws + M.off_cstart);
    b.cell_cur = (int*)(ws + M.off_ccur);
    b.chunks = (int2*)(ws + M.off_chunks);
    b.n_chunks_img = (int*)(ws + M.off_nchunks);
    b.best = (int*)(ws + M.off_best);
    b.b1 = (double*)(ws + M.off_b1);
    b.qcell = (int*)(ws + M.off_qcell);
    b.kth = (double*)(ws + M.off_kth);
    b.cbest = (int*)(ws + M.off_cbest);
    b.cb12 = (float*)(ws + M.off_cb12);
    b.tkey = (unsigned long long*)(ws + M.off_tkey);
    b.counters = (int*)(ws + M.off_cnt);
    b.out_q = (int*)(ws + M.off_outq);
    b.out_t = (int*)(ws + M.off_outt);
    b.out_n = (int*)(ws + M.off_outn);
    return b;
}


// launch_info: [0] tag, [1] n_img, [2] n_pair, [3] queries, [4] largest query set, [5] setup
// blocks, [6] kc, [7] workspace address, [8..] the MatchBatch record, [last] MFMA chunk rows.
static void fill_launch_info(const MatchPlanHost& M, const MatchBatch& b, int kc, const void* ws, long long* info) {
    const int n_pair = (int)M.pair.size();
    int max_na = 0, max_chunks = 0;
    for (int p = 0; p < n_pair; p++) {
        max_na = std::max(max_na, M.img[M.pair[p].a].n);
        max_chunks = std::max(max_chunks, M.img[M.pair[p].a].chunk_cap);
    }
    const long long extra =
        std::min<long long>(std::max<long long>(1, (std::max(M.n_targets, M.n_kp) + SU_THREADS - 1) / SU_THREADS), 4096);
    memset(info, 0, 8 * FSX_MATCH_INFO_LEN);
    info[0] = 0x6673786d;     // tag checked by fsx_match_launch
    info[1] = (long long)M.img.size();
    info[2] = n_pair;
    info[3] = M.n_query;
    info[4] = max_na;
    info[5] = (long long)M.img.size() + extra;
    info[6] = kc;
    info[7] = (long long)(intptr_t)ws;
    info[FSX_MATCH_INFO_LEN - 1] = max_chunks;
    memcpy(info + 8, &b, sizeof(b));
}

// The matcher's launches (setup, kc-nearest prefilter, contraction, compaction) on `st`.
// contract_only: stop after the contraction, which then publishes nothing (the learned
// filter's stage runs k_m_publish and the compaction afterwards).
static int launch_sequence(const long long* info, int form, hipEvent_t* ev, hipStream_t st,
                           bool contract_only = false) {
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    b.defer = contract_only ? 1 : 0;
    const long long nq = info[3];
    const int n_pair = (int)info[2], max_na = (int)info[4];
    if (ev) FSX_HIP(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(k_m_setup, dim3((unsigned)info[5]), dim3(SU_THREADS), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (nq > 0 && info[6] > 0) {
        const dim3 g((unsigned)((max_na + 255) / 256), n_pair);
        const long long kc = info[6];
        hipLaunchKernelGGL(k_m_knn, g, dim3(256), 0, st, b);
        if (kc <= 8) hipLaunchKernelGGL(k_m_knn_exact<8>, g, dim3(256), 0, st, b);
        else if (kc <= 16) hipLaunchKernelGGL(k_m_knn_exact<16>, g, dim3(256), 0, st, b);
        else if (kc <= 24) hipLaunchKernelGGL(k_m_knn_exact<24>, g, dim3(256), 0, st, b);
        else if (kc <= 32) hipLaunchKernelGGL(k_m_knn_exact<32>, g, dim3(256), 0, st, b);
        else if (kc <= 48) hipLaunchKernelGGL(k_m_knn_exact<48>, g, dim3(256), 0, st, b);
        else hipLaunchKernelGGL(k_m_knn_exact<KMAX>, g, dim3(256), 0, st, b);
    }
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[1], st));
    if (nq > 0) {
        if (form == 0)
            hipLaunchKernelGGL(k_m_plain, dim3((unsigned)((max_na + 255) / 256), n_pair), dim3(256), 0, st, b);
        else
        {
            // the contraction's workgroups: one per (chunk, pair), or a fixed number of
            // persistent ones striding over the items (TUNE_MATCH_BLOCKS)
            const int mc = (int)info[FSX_MATCH_INFO_LEN - 1], pers = tune_get(TUNE_MATCH_BLOCKS);
            const long long items = (long long)mc * n_pair;
            if (pers > 0 && items > 0 && items < (1ll << 31)) {
                hipLaunchKernelGGL(k_m_mfma<true>, dim3((unsigned)std::min<long long>(pers, items)), dim3(64), 0, st, b,
                                   mc);
            } else {
                hipLaunchKernelGGL(k_m_mfma<false>, dim3((unsigned)std::max(mc, 1), n_pair), dim3(64), 0, st, b, 0);
            }
        }
        FSX_LAUNCH_CHECK();
    }
    if (ev) FSX_HIP(hipEventRecord(ev[2], st));
    if (contract_only) return FSX_OK;
    hipLaunchKernelGGL(k_m_compact, dim3((unsigned)n_pair), dim3(CP_THREADS), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[3], st));
    return FSX_OK;
}


// ---- matcher without allocation ---------------------------------------------------------
// fsx_match_batch_noalloc keeps its workspace in this file-scope array (one copy per GPU,
// held until the process ends) and writes its results straight into the caller's buffers.
// The host never uses the array's name as an address: scratch_base() looks the device copy
// up once per device with hipGetSymbolAddress and keeps that pointer.
constexpr size_t MATCH_SCRATCH_BYTES = (size_t)32 << 20;
constexpr int NOALLOC_MAX_IMG = 64;
constexpr int NOALLOC_MAX_PAIR = 256;
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[MATCH_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[MAX_DEVICES];

static int scratch_base(void** out) {
    int dev = -1;
    FSX_HIP(hipGetDevice(&dev));
    FSX_REQUIRE(dev >= 0 && dev < MAX_DEVICES, "device %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        FSX_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        FSX_REQUIRE(p != nullptr, "no device address for the matcher scratch");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FSX_OK;
}

__global__ void k_m_zero_int(int* p) { *p = 0; }

static int check_info(const void* ws, const long long* info) {
    FSX_REQUIRE(ws && info, "null pointer argument");
    FSX_REQUIRE(info[0] == 0x6673786d, "launch_info was not filled by fsx_match_prepare");
    FSX_REQUIRE(info[7] == (long long)(intptr_t)ws, "launch_info belongs to another workspace");
    return FSX_OK;
}

// Host tables reach the device as kernel arguments (no staging buffer, no host wait): each
// launch of k_put_bytes stores up to PUT_CHUNK bytes of its argument at base + off.
constexpr int PUT_CHUNK = 2016;
struct PutBytes {
    unsigned long long off;
    int n, pad;
    unsigned char b[PUT_CHUNK];
};

__global__ __launch_bounds__(256) void k_put_bytes(unsigned char* base, const PutBytes P) {
    for (int i = threadIdx.x; i < P.n; i += blockDim.x) base[P.off + i] = P.b[i];
}

static int put_bytes(unsigned char* base, size_t off, const void* src, size_t n, hipStream_t st) {
    PutBytes P;
    memset(&P, 0, sizeof(P));
    for (size_t o = 0; o < n; o += PUT_CHUNK) {
        const size_t k = std::min(n - o, (size_t)PUT_CHUNK);
        P.off = off + o;
        P.n = (int)k;
        memcpy(P.b, (const char*)src + o, k);
        hipLaunchKernelGGL(k_put_bytes, dim3(1), dim3(256), 0, st, base, P);
        FSX_LAUNCH_CHECK();
    }
    return FSX_OK;
}

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

int fsx_match_prepare(int n_img, const int* n_kp, const int* W, const int* H, const uint8_t* desc_dev,
                      const double* xy_dev, int n_pair, const int* pa, const int* pb, const int* dense,
                      const double* Hm, double radius, int kc, double ratio, void* ws, size_t ws_bytes,
                      long long* out_offsets, long long* launch_info, void* stream) {
    FSX_REQUIRE(desc_dev && xy_dev && ws && pa && pb && launch_info, "null pointer argument");
    FSX_REQUIRE(kc >= 0 && kc <= KMAX, "kc must be in [0, %d]", KMAX);
    static_assert(sizeof(MatchBatch) <= 8 * (FSX_MATCH_INFO_LEN - 9), "launch_info too small for MatchBatch");
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, dense, Hm, radius, M);
    if (rc) return rc;
    FSX_REQUIRE(ws_bytes >= M.total, "match workspace too small: %zu < %zu", ws_bytes, M.total);
    FSX_REQUIRE(n_img <= (1 << 20) && n_pair <= 65535, "too many images or pairs in one batch");
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)ws;
    MatchBatch b = batch_ptrs(M, w, desc_dev, xy_dev, radius, kc, ratio);
    rc = upload_async(w + M.off_img, M.img.data(), sizeof(MImage) * M.img.size(), st);
    if (rc) return rc;
    rc = upload_async(w + M.off_pair, M.pair.data(), sizeof(MPair) * M.pair.size(), st);
    if (rc) return rc;
    if (out_offsets)
        for (int p = 0; p < n_pair; p++) out_offsets[p] = M.pair[p].o_off;
    fill_launch_info(M, b, kc, ws, launch_info);
    return FSX_OK;
}

int fsx_match_launch(void* ws, const long long* info, int form, void* stage_events_v, void* stream) {
    FSX_REQUIRE(ws && info, "null pointer argument");
    FSX_REQUIRE(info[0] == 0x6673786d, "launch_info was not filled by fsx_match_prepare");
    FSX_REQUIRE(form == 0 || form == 1, "form must be 0 (plain FP32) or 1 (MFMA)");
    FSX_REQUIRE(info[7] == (long long)(intptr_t)ws, "launch_info belongs to another workspace");
    return launch_sequence(info, form, (hipEvent_t*)stage_events_v, (hipStream_t)stream);
}

int fsx_match_batch(int n_img, const int* n_kp, const int* W, const int* H, const uint8_t* desc_dev,
                    const double* xy_dev, int n_pair, const int* pa, const int* pb, const int* dense,
                    const double* Hm, double radius, int kc, double ratio, int form, void* ws, size_t ws_bytes,
                    long long* out_offsets, void* stage_events_v, void* stream) {
    FSX_REQUIRE(form == 0 || form == 1, "form must be 0 (plain FP32) or 1 (MFMA)");
    long long info[FSX_MATCH_INFO_LEN];
    int rc = fsx_match_prepare(n_img, n_kp, W, H, desc_dev, xy_dev, n_pair, pa, pb, dense, Hm, radius, kc, ratio, ws,
                               ws_bytes, out_offsets, info, stream);
    if (rc) return rc;
    return fsx_match_launch(ws, info, form, stage_events_v, stream);
}

int fsx_match_filter_features(void* ws, const long long* info, const float* size_dev, const float* angle_dev,
                              void* X_dev, void* stream) {
    FSX_REQUIRE(ws && info && size_dev && angle_dev && X_dev, "null pointer argument");
    FSX_REQUIRE(info[0] == 0x6673786d && info[7] == (long long)(intptr_t)ws, "launch_info does not match ws");
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    const int n_pair = (int)info[2], max_na = (int)info[4];
    if (info[3] == 0) return FSX_OK;
    hipLaunchKernelGGL(k_m_features, dim3((unsigned)((max_na + 3) / 4), n_pair), dim3(256), 0, (hipStream_t)stream, b,
                       size_dev, angle_dev, (unsigned short*)X_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Copies the results of the last fsx_match_batch on `ws` to host arrays (synchronises):
// counts[n_pair], then q/t of all pairs at the offsets fsx_match_batch reported.
int fsx_match_results(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                      const int* pb, double radius, const void* ws, int* counts, int* q, int* t, int* flags,
                      void* stream) {
    FSX_REQUIRE(ws && (counts || n_pair <= 0), "null pointer argument");
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

// The last batch's (q, t) lists packed on the device (asynchronous): pair p's counts[p]
// entries (the counts of fsx_match_results) to rows poff_dev[p] .. of q_dev / t_dev (int32,
// cap rows each); poff_dev: n_pair int64 row offsets, normally the exclusive prefix sums of the
// counts.  Only the matches cross to the host afterwards, not the per-query output slots.
int fsx_match_pack(const void* ws, const long long* info, const long long* poff_dev, long long cap, int* q_dev,
                   int* t_dev, void* stream) {
    FSX_REQUIRE(ws && info && poff_dev && q_dev && t_dev, "null pointer argument");
    FSX_REQUIRE(info[0] == 0x6673786d && info[7] == (long long)(intptr_t)ws, "launch_info does not match ws");
    FSX_REQUIRE(cap >= 0, "negative capacity");
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    const int n_pair = (int)info[2];
    if (n_pair <= 0) return FSX_OK;
    hipLaunchKernelGGL(k_m_pack, dim3((unsigned)n_pair), dim3(256), 0, (hipStream_t)stream, b, poff_dev, cap, q_dev,
                       t_dev);
    FSX_LAUNCH_CHECK();
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

// The INT8 form's assumption: for v_mfma_i32_32x32x32_i8, lane l supplies row l % 32 of A and
// column l % 32 of B, both with the same 16 k positions for its half h = l / 32, and the
// result layout is the one of the 32x32 BF16 instruction.  Operands: A byte e of lane l =
// ((l%32)*7 + h*3 + e*5) % 11 - 5, B byte = ((l%32)*3 + h*5 + e*7) % 13 - 6; out: 64 x 16 int32.
__global__ void k_mfma_i8_layout_probe(int* out) {
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    i32x4_t a, b;
#pragma unroll
    for (int w = 0; w < 4; w++) {
        unsigned wa = 0, wb = 0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int e = 4 * w + k;
            const int av = (r * 7 + h * 3 + e * 5) % 11 - 5;
            const int bv = (r * 3 + h * 5 + e * 7) % 13 - 6;
            wa |= ((unsigned)av & 0xffu) << (8 * k);
            wb |= ((unsigned)bv & 0xffu) << (8 * k);
        }
        a[w] = (int)wa;
        b[w] = (int)wb;
    }
    i32x16_t acc = {};
    acc = __builtin_amdgcn_mfma_i32_32x32x32_i8(a, b, acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; i++) out[lane * 16 + i] = acc[i];
}

int fsx_mfma_i8_layout_probe(int* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_mfma_i8_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_mfma_layout_probe(float* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_mfma_layout_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Byte offsets inside the workspace of the per-query candidate arrays of the last batch (in
// the caller's query order): [0] cbest (int32, best candidate before the ratio test, -1 = none), [1] cb12 (float32 pairs: its distance, the second smallest), [2] best
// (int32, accepted target or -1), [3] qcell (int32).  For the learned filter.
int fsx_match_candidate_offsets(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                                const int* pb, double radius, long long* offs) {
    FSX_REQUIRE(offs, "null pointer argument");
    MatchPlanHost M;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, nullptr, nullptr, radius, M);
    if (rc) return rc;
    offs[0] = (long long)M.off_cbest;
    offs[1] = (long long)M.off_cb12;
    offs[2] = (long long)M.off_best;
    offs[3] = (long long)M.off_qcell;
    return FSX_OK;
}

// Per-query state of the last batch (synchronises): accepted best target (-1 = none), its
// distance b1, and the query's grid cell (-1 = predicted outside).  For tests and diagnostics.
int fsx_match_query_state(int n_img, const int* n_kp, const int* W, const int* H, int n_pair, const int* pa,
                          const int* pb, double radius, const void* ws, int* best, double* b1, int* qcell,
                          void* stream) {
    FSX_REQUIRE(ws, "null pointer argument");
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


// fsx_match_batch with no allocation of any kind (see fsx.h): plans into stack tables, keeps
// its workspace in the per-device scratch array, and points the compaction's outputs and the
// flag words at the caller's buffers.
int fsx_match_batch_noalloc(int n_img, const int* n_kp, const int* W, const int* H, const uint8_t* desc_dev,
                            const double* xy_dev, int n_pair, const int* pa, const int* pb, const int* dense,
                            const double* Hm, double radius, int kc, double ratio, int form, int* q_dev, int* t_dev,
                            int* counts_dev, int* flags_dev, long long* out_offsets, void* stream) {
    FSX_REQUIRE(desc_dev && xy_dev && n_kp && W && H && pa && pb && q_dev && t_dev && counts_dev && flags_dev,
                "null pointer argument");
    FSX_REQUIRE(form == 0 || form == 1, "form must be 0 (plain FP32) or 1 (MFMA)");
    FSX_REQUIRE(kc >= 0 && kc <= KMAX, "kc must be in [0, %d]", KMAX);
    FSX_REQUIRE(n_img >= 1 && n_img <= NOALLOC_MAX_IMG, "n_img must be in [1, %d] here", NOALLOC_MAX_IMG);
    FSX_REQUIRE(n_pair >= 1 && n_pair <= NOALLOC_MAX_PAIR, "n_pair must be in [1, %d] here", NOALLOC_MAX_PAIR);
    MImage img_tab[NOALLOC_MAX_IMG];
    MPair pair_tab[NOALLOC_MAX_PAIR];
    MatchPlanHost M;
    M.img.ext = img_tab;
    M.img.cap = NOALLOC_MAX_IMG;
    M.pair.ext = pair_tab;
    M.pair.cap = NOALLOC_MAX_PAIR;
    int rc = match_plan(n_img, n_kp, W, H, n_pair, pa, pb, dense, Hm, radius, M);
    if (rc) return rc;
    // everything below the flag words lives in scratch; flags and outputs are the caller's
    FSX_REQUIRE(M.off_cnt <= MATCH_SCRATCH_BYTES,
                "batch needs %zu bytes of matcher scratch, above the %zu this entry point holds "
                "(fsx_match_batch takes a workspace of any size)", M.off_cnt, MATCH_SCRATCH_BYTES);
    void* base = nullptr;
    rc = scratch_base(&base);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    unsigned char* w = (unsigned char*)base;
    MatchBatch b = batch_ptrs(M, (char*)w, desc_dev, xy_dev, radius, kc, ratio);
    b.counters = flags_dev;
    b.out_q = q_dev;
    b.out_t = t_dev;
    b.out_n = counts_dev;
    rc = put_bytes(w, M.off_img, img_tab, sizeof(MImage) * (size_t)n_img, st);
    if (rc) return rc;
    rc = put_bytes(w, M.off_pair, pair_tab, sizeof(MPair) * (size_t)n_pair, st);
    if (rc) return rc;
    if (out_offsets)
        for (int p = 0; p < n_pair; p++) out_offsets[p] = M.pair[p].o_off;
    long long info[FSX_MATCH_INFO_LEN];
    fill_launch_info(M, b, kc, base, info);
    return launch_sequence(info, form, nullptr, st);
}


// ---- learned-filter stage: contraction, accepted list, filter rows, filtered publication ----

int fsx_match_contract(void* ws, const long long* info, int form, void* stream) {
    int rc = check_info(ws, info);
    if (rc) return rc;
    FSX_REQUIRE(form == 0 || form == 1, "form must be 0 (plain FP32) or 1 (MFMA)");
    return launch_sequence(info, form, nullptr, (hipStream_t)stream, true);
}

int fsx_match_accepted(void* ws, const long long* info, int* list_dev, int* count_dev, void* stream) {
    int rc = check_info(ws, info);
    if (rc) return rc;
    FSX_REQUIRE(list_dev && count_dev, "null pointer argument");
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_m_zero_int, dim3(1), dim3(1), 0, st, count_dev);
    FSX_LAUNCH_CHECK();
    const int n_pair = (int)info[2], max_na = (int)info[4];
    if (info[3] == 0 || max_na == 0) return FSX_OK;
    hipLaunchKernelGGL(k_m_accepted, dim3((unsigned)((max_na + 255) / 256), n_pair), dim3(256), 0, st, b,
                       (int2*)list_dev, count_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_match_filter_rows(void* ws, const long long* info, const float* size_dev, const float* angle_dev,
                          const int* list_dev, const int* count_dev, long long cap, void* X_dev, void* stream) {
    int rc = check_info(ws, info);
    if (rc) return rc;
    FSX_REQUIRE(size_dev && angle_dev && list_dev && count_dev && X_dev, "null pointer argument");
    FSX_REQUIRE(cap >= 0 && cap <= info[3], "row capacity %lld outside [0, %lld]", cap, info[3]);
    if (cap == 0) return FSX_OK;
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    hipLaunchKernelGGL(k_m_features_list, dim3((unsigned)((cap + 3) / 4)), dim3(256), 0, (hipStream_t)stream, b,
                       size_dev, angle_dev, (const int2*)list_dev, count_dev, (unsigned short*)X_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_match_publish(void* ws, const long long* info, const int* list_dev, const int* count_dev, long long cap,
                      const float* score_dev, float tau, void* stream) {
    int rc = check_info(ws, info);
    if (rc) return rc;
    FSX_REQUIRE(list_dev && count_dev, "null pointer argument");
    FSX_REQUIRE(cap >= 0 && cap <= info[3], "list capacity %lld outside [0, %lld]", cap, info[3]);
    MatchBatch b;
    memcpy(&b, info + 8, sizeof(b));
    hipStream_t st = (hipStream_t)stream;
    if (cap > 0) {
        hipLaunchKernelGGL(k_m_publish, dim3((unsigned)((cap + 255) / 256)), dim3(256), 0, st, b,
                           (const int2*)list_dev, count_dev, score_dev, tau);
        FSX_LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(k_m_compact, dim3((unsigned)info[2]), dim3(CP_THREADS), 0, st, b);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
