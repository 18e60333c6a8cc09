// Geometric verification on gfx950: a deterministic homography RANSAC per image pair (the role
// of matching.py:226 cv2.findHomography(..., RANSAC) in match_pair; cv2's own sampling is not
// reproducible, this one is).  It labels the plain pipeline's matches as inliers for the
// learned match filter.
//
// Per pair, hypothesis h = 0 .. n_hyp-1 draws 4 distinct matches with a counter hash
// (splitmix64 of (seed, pair, h, draw)), solves the 8x8 DLT system (h33 = 1) by Gaussian
// elimination with partial pivoting in float64 and counts matches whose reprojection error is
// below thr; the hypothesis with the most inliers wins (ties: lowest h), and its inliers are
// the labels.  Plain float64 arithmetic without contraction, in a fixed order, so the CPU
// restatement reproduces every count.
#include "fsx_common.h"
#include <math.h>

namespace fsx {

__host__ __device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ unsigned long long draw(unsigned long long seed, int p, int h, int j) {
    return splitmix64(seed ^ splitmix64(((unsigned long long)p << 40) ^ ((unsigned long long)h << 8) ^ (unsigned)j));
}

// solve the 4-point DLT; returns false when a pivot is below 1e-12 (degenerate sample)
__device__ bool solve_h4(const double* xa, const double* ya, const double* xb, const double* yb, double* h) {
    double A[8][9];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const double x = xa[k], y = ya[k], u = xb[k], v = yb[k];
        double* r0 = A[2 * k];
        double* r1 = A[2 * k + 1];
        r0[0] = x; r0[1] = y; r0[2] = 1.0; r0[3] = 0.0; r0[4] = 0.0; r0[5] = 0.0;
        r0[6] = -(u * x); r0[7] = -(u * y); r0[8] = u;
        r1[0] = 0.0; r1[1] = 0.0; r1[2] = 0.0; r1[3] = x; r1[4] = y; r1[5] = 1.0;
        r1[6] = -(v * x); r1[7] = -(v * y); r1[8] = v;
    }
    for (int c = 0; c < 8; c++) {
        int piv = c;
        double best = fabs(A[c][c]);
        for (int r = c + 1; r < 8; r++) {
            const double a = fabs(A[r][c]);
            if (a > best) {
                best = a;
                piv = r;
            }
        }
        if (best < 1e-12) return false;
        if (piv != c)
            for (int k = 0; k < 9; k++) {
                const double t = A[c][k];
                A[c][k] = A[piv][k];
                A[piv][k] = t;
            }
        for (int r = c + 1; r < 8; r++) {
            const double f = A[r][c] / A[c][c];
            for (int k = c; k < 9; k++) A[r][k] = A[r][k] - f * A[c][k];
        }
    }
    for (int c = 7; c >= 0; c--) {
        double s = A[c][8];
        for (int k = c + 1; k < 8; k++) s = s - A[c][k] * h[k];
        h[c] = s / A[c][c];
    }
    h[8] = 1.0;
    return true;
}

__device__ __forceinline__ bool is_inlier(const double* h, double x, double y, double u, double v, double thr2) {
    const double w = (h[6] * x + h[7] * y) + h[8];
    if (!(fabs(w) > 1e-12)) return false;
    const double px = ((h[0] * x + h[1] * y) + h[2]) / w;
    const double py = ((h[3] * x + h[4] * y) + h[5]) / w;
    const double dx = px - u, dy = py - v;
    return dx * dx + dy * dy < thr2;
}

__global__ __launch_bounds__(256) void k_ransac(const long long* __restrict__ off, const double* __restrict__ pa,
                                                const double* __restrict__ pb, int n_hyp, double thr2,
                                                unsigned long long seed, uint8_t* __restrict__ inlier,
                                                double* __restrict__ H_out, int* __restrict__ n_in) {
    __shared__ int s_cnt[256], s_hyp[256];
    __shared__ double s_h[9];
    const int p = blockIdx.x, tid = threadIdx.x;
    const long long o = off[p];
    const int n = (int)(off[p + 1] - o);
    const double* A = pa + 2 * o;
    const double* Bq = pb + 2 * o;
    int best = -1, best_h = 0x7fffffff;
    if (n >= 4) {
        for (int hy = tid; hy < n_hyp; hy += 256) {
            int idx[4];
            bool ok = true;
            for (int j = 0; j < 4 && ok; j++) {
                int cand = -1;
                for (int attempt = 0; attempt < 16; attempt++) {
                    const int c = (int)(draw(seed, p, hy, j * 16 + attempt) % (unsigned long long)n);
                    bool dup = false;
                    for (int u = 0; u < j; u++) dup |= idx[u] == c;
                    if (!dup) {
                        cand = c;
                        break;
                    }
                }
                if (cand < 0) ok = false;
                else idx[j] = cand;
            }
            if (!ok) continue;
            double xa[4], ya[4], xb[4], yb[4], h[9];
            for (int j = 0; j < 4; j++) {
                xa[j] = A[2 * idx[j]];
                ya[j] = A[2 * idx[j] + 1];
                xb[j] = Bq[2 * idx[j]];
                yb[j] = Bq[2 * idx[j] + 1];
            }
            if (!solve_h4(xa, ya, xb, yb, h)) continue;
            int cnt = 0;
            for (int i = 0; i < n; i++) cnt += is_inlier(h, A[2 * i], A[2 * i + 1], Bq[2 * i], Bq[2 * i + 1], thr2);
            if (cnt > best || (cnt == best && hy < best_h)) {
                best = cnt;
                best_h = hy;
            }
        }
    }
    s_cnt[tid] = best;
    s_hyp[tid] = best_h;
    __syncthreads();
    if (tid == 0) {
        int bc = -1, bh = 0x7fffffff;
        for (int u = 0; u < 256; u++)
            if (s_cnt[u] > bc || (s_cnt[u] == bc && s_hyp[u] < bh)) {
                bc = s_cnt[u];
                bh = s_hyp[u];
            }
        // rebuild the winning model (the same draws and solve as above)
        bool have = false;
        if (bc >= 0) {
            int idx[4];
            bool ok = true;
            for (int j = 0; j < 4 && ok; j++) {
                int cand = -1;
                for (int attempt = 0; attempt < 16; attempt++) {
                    const int c = (int)(draw(seed, p, bh, j * 16 + attempt) % (unsigned long long)n);
                    bool dup = false;
                    for (int u = 0; u < j; u++) dup |= idx[u] == c;
                    if (!dup) {
                        cand = c;
                        break;
                    }
                }
                if (cand < 0) ok = false;
                else idx[j] = cand;
            }
            double xa[4], ya[4], xb[4], yb[4], h[9];
            for (int j = 0; ok && j < 4; j++) {
                xa[j] = A[2 * idx[j]];
                ya[j] = A[2 * idx[j] + 1];
                xb[j] = Bq[2 * idx[j]];
                yb[j] = Bq[2 * idx[j] + 1];
            }
            have = ok && solve_h4(xa, ya, xb, yb, h);
            for (int k = 0; k < 9; k++) s_h[k] = have ? h[k] : 0.0;
        }
        if (!have)
            for (int k = 0; k < 9; k++) s_h[k] = 0.0;
        n_in[p] = have ? bc : 0;
        for (int k = 0; k < 9; k++) H_out[9 * p + k] = s_h[k];
    }
    __syncthreads();
    const bool have = s_h[8] == 1.0;
    double h[9];
    for (int k = 0; k < 9; k++) h[k] = s_h[k];
    for (int i = tid; i < n; i += 256)
        inlier[o + i] = (uint8_t)(have && is_inlier(h, A[2 * i], A[2 * i + 1], Bq[2 * i], Bq[2 * i + 1], thr2));
}

}  // namespace fsx

using namespace fsx;

extern "C" {

// Deterministic homography RANSAC for n_pair sets of matches: pair p's matches are rows
// off_dev[p] .. off_dev[p+1] of pa_dev / pb_dev (float64 (x, y) in image a / image b).  n_hyp
// hypotheses per pair, inlier iff reprojection error < thr px.  Writes inlier_dev (one byte per
// match), H_dev (9 float64 per pair, all zero when no model) and n_in_dev (per pair).
int fsx_ransac_homography(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev, int n_hyp,
                          double thr, unsigned long long seed, uint8_t* inlier_dev, double* H_dev, int* n_in_dev,
                          void* stream) {
    FSX_REQUIRE(n_pair >= 0 && n_hyp >= 1 && n_hyp <= (1 << 20) && thr > 0, "bad RANSAC arguments");
    if (n_pair == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && pa_dev && pb_dev && inlier_dev && H_dev && n_in_dev, "null pointer argument");
    hipLaunchKernelGGL(k_ransac, dim3(n_pair), dim3(256), 0, (hipStream_t)stream, off_dev, pa_dev, pb_dev, n_hyp,
                       thr * thr, seed, inlier_dev, H_dev, n_in_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
