// Geometric verification on gfx950: a deterministic homography RANSAC per image pair (the role
// of matching.py:226 cv2.findHomography(..., RANSAC) in match_pair; cv2's own sampling is not
// reproducible, this one is).  It labels the plain pipeline's matches as inliers for the
// learned match filter.
//
// Per pair, hypothesis h = 0 .. n_hyp-1 draws 4 distinct matches with a counter hash
// (splitmix64 of (seed, pair, h, draw)), solves the 8x8 DLT system (h33 = 1) by Gaussian
// elimination with partial pivoting in float64 and counts matches whose reprojection error is
// below thr; the hypothesis with the most inliers wins (ties: lowest h), and its inliers are
// the labels; the winning homography is then refit on them.  The same scheme serves the
// fundamental matrix (8 points, rank 2) and the similarity (2 points).  Float64 arithmetic in a
// fixed order: the solves without contraction, the inlier tests with explicit fused
// multiply-adds (fma, one rounding each), so the CPU restatement (plain operations and C's fma)
// reproduces every count and model.
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

// Step C of the partial-pivot elimination of the 8 x 9 DLT system: the pivot is the largest
// |A[r][C]| of rows C.. (the first of equal ones), swapped to row C; the rows below are reduced.
// Compile-time indices only (the swap is a select), so the system stays in registers; only
// columns >= C are swapped, the others are not read again.  false: pivot below 1e-12.
template <int C>
__device__ __forceinline__ bool h4_step(double (&A)[8][9]) {
    int piv = C;
    double best = fabs(A[C][C]);
#pragma unroll
    for (int r = C + 1; r < 8; r++) {
        const double a = fabs(A[r][C]);
        const bool gt = a > best;
        best = gt ? a : best;
        piv = gt ? r : piv;
    }
    if (best < 1e-12) return false;
#pragma unroll
    for (int r = C + 1; r < 8; r++) {
        const bool sw = piv == r;
#pragma unroll
        for (int k = C; k < 9; k++) {
            const double t = A[C][k];
            A[C][k] = sw ? A[r][k] : t;
            A[r][k] = sw ? t : A[r][k];
        }
    }
#pragma unroll
    for (int r = C + 1; r < 8; r++) {
        const double f = A[r][C] / A[C][C];
#pragma unroll
        for (int k = C; k < 9; k++) A[r][k] = A[r][k] - f * A[C][k];
    }
    return true;
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
    if (!h4_step<0>(A) || !h4_step<1>(A) || !h4_step<2>(A) || !h4_step<3>(A) || !h4_step<4>(A) || !h4_step<5>(A) ||
        !h4_step<6>(A) || !h4_step<7>(A))
        return false;
#pragma unroll
    for (int c = 7; c >= 0; c--) {
        double s = A[c][8];
#pragma unroll
        for (int k = c + 1; k < 8; k++) s = s - A[c][k] * h[k];
        h[c] = s / A[c][c];
    }
    h[8] = 1.0;
    return true;
}

// reprojection error of (x, y) -> (u, v) below sqrt(thr2), without a division:
// |H p / w - q|^2 < thr2  <=>  |H p - w q|^2 < thr2 w^2 (w != 0); fused multiply-adds in a
// fixed nesting (the CPU restatement uses C's fma, so the bits agree), no branch
__device__ __forceinline__ bool is_inlier(const double* h, double x, double y, double u, double v, double thr2) {
    const double w = fma(h[6], x, fma(h[7], y, h[8]));
    const double ex = fma(-u, w, fma(h[0], x, fma(h[1], y, h[2])));
    const double ey = fma(-v, w, fma(h[3], x, fma(h[4], y, h[5])));
    return ((int)(fabs(w) > 1e-12) & (int)(fma(ex, ex, ey * ey) < thr2 * (w * w))) != 0;   // both tests, no branch
}

// ---- models: K points per sample, solve(K points) -> 9 parameters, inlier test -------------

// Homography (the role of cv2.findHomography in match_pair, matching.py:226): 4-point DLT,
// h33 = 1, reprojection error in image b.
struct ModelH {
    static constexpr int K = 4;
    static constexpr bool refit = true;
    __device__ static bool solve(const double* xa, const double* ya, const double* xb, const double* yb, double* m) {
        return solve_h4(xa, ya, xb, yb, m);
    }
    __device__ static bool inlier(const double* m, double x, double y, double u, double v, double thr2) {
        return is_inlier(m, x, y, u, v, thr2);
    }
};

// The closest rank-2 matrix (Frobenius norm) to the 3 x 3 row-major f, in place: f is scaled
// by its largest |entry|, then replaced by f (I - v v^T), v the eigenvector of f^T f with the
// smallest eigenvalue (f v v^T = sigma3 u3 v3^T, the last term of f's SVD).  v comes from six
// cyclic Jacobi sweeps over f^T f (rotations (0,1), (0,2), (1,2); tan of the angle from
// theta = (a_qq - a_pp) / (2 a_pq), the smaller root), ties of the smallest diagonal to the
// lower index.  Only +, -, *, / and sqrt in a fixed order, so a CPU restatement gives the same
// bits.
__device__ void rank2_project(double* f) {
    double mx = 0.0;
    for (int k = 0; k < 9; k++) mx = fmax(mx, fabs(f[k]));
    if (!(mx > 0.0)) return;
    for (int k = 0; k < 9; k++) f[k] = f[k] / mx;
    double S[3][3], V[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            S[i][j] = (f[i] * f[j] + f[3 + i] * f[3 + j]) + f[6 + i] * f[6 + j];
            V[i][j] = i == j ? 1.0 : 0.0;
        }
    for (int sweep = 0; sweep < 6; sweep++)
        for (int r = 0; r < 3; r++) {
            const int p = r == 2 ? 1 : 0, q = r == 0 ? 1 : 2;
            const double apq = S[p][q];
            if (apq == 0.0) continue;
            const double theta = (S[q][q] - S[p][p]) / (2.0 * apq);
            const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
            const double c = 1.0 / sqrt(t * t + 1.0), sn = t * c;
            // S <- R^T S R and V <- V R, R = I but R[p][p] = R[q][q] = c, R[p][q] = sn, R[q][p] = -sn
            for (int k = 0; k < 3; k++) {
                const double skp = S[k][p], skq = S[k][q];
                S[k][p] = c * skp - sn * skq;
                S[k][q] = sn * skp + c * skq;
                const double vkp = V[k][p], vkq = V[k][q];
                V[k][p] = c * vkp - sn * vkq;
                V[k][q] = sn * vkp + c * vkq;
            }
            for (int k = 0; k < 3; k++) {
                const double spk = S[p][k], sqk = S[q][k];
                S[p][k] = c * spk - sn * sqk;
                S[q][k] = sn * spk + c * sqk;
            }
        }
    // kmin: the smallest diagonal entry, ties to the lower index (selects, no indexed access)
    const bool k1 = S[1][1] < S[0][0];
    const double s01 = k1 ? S[1][1] : S[0][0];
    const bool k2 = S[2][2] < s01;
    const double v0 = k2 ? V[0][2] : (k1 ? V[0][1] : V[0][0]);
    const double v1 = k2 ? V[1][2] : (k1 ? V[1][1] : V[1][0]);
    const double v2 = k2 ? V[2][2] : (k1 ? V[2][1] : V[2][0]);
    for (int i = 0; i < 3; i++) {
        const double w = (f[3 * i] * v0 + f[3 * i + 1] * v1) + f[3 * i + 2] * v2;
        f[3 * i] = f[3 * i] - w * v0;
        f[3 * i + 1] = f[3 * i + 1] - w * v1;
        f[3 * i + 2] = f[3 * i + 2] - w * v2;
    }
}

// Step C of the full-pivot elimination of the 8 x 9 system: the pivot is the largest |entry|
// of rows C.. and columns C.. (first in row-major order), swapped to (C, C); the rows below are
// reduced.  Every index is a compile-time constant (the swaps and the column permutation are
// selects), so the system stays in registers instead of scratch memory.  Only the entries read
// later are swapped: columns >= C of the pivot rows.  false: pivot below 1e-12 (degenerate).
template <int C>
__device__ __forceinline__ bool f8_step(double (&A)[8][9], int (&perm)[9]) {
    int pr = C, pc = C;
    double best = -1.0;
#pragma unroll
    for (int r = C; r < 8; r++)
#pragma unroll
        for (int k = C; k < 9; k++) {
            const double a = fabs(A[r][k]);
            const bool gt = a > best;
            best = gt ? a : best;
            pr = gt ? r : pr;
            pc = gt ? k : pc;
        }
    if (best < 1e-12) return false;
#pragma unroll
    for (int r = C + 1; r < 8; r++) {
        const bool sw = pr == r;
#pragma unroll
        for (int k = C; k < 9; k++) {
            const double t = A[C][k];
            A[C][k] = sw ? A[r][k] : t;
            A[r][k] = sw ? t : A[r][k];
        }
    }
#pragma unroll
    for (int k = C + 1; k < 9; k++) {
        const bool sw = pc == k;
#pragma unroll
        for (int r = 0; r < 8; r++) {
            const double t = A[r][C];
            A[r][C] = sw ? A[r][k] : t;
            A[r][k] = sw ? t : A[r][k];
        }
        const int t = perm[C];
        perm[C] = sw ? perm[k] : t;
        perm[k] = sw ? t : perm[k];
    }
#pragma unroll
    for (int r = C + 1; r < 8; r++) {
        const double f = A[r][C] / A[C][C];
#pragma unroll
        for (int k = C; k < 9; k++) A[r][k] = A[r][k] - f * A[C][k];
    }
    return true;
}

// Fundamental matrix (the role of cv2.findFundamentalMat in _verify, matching.py:175-191):
// normalized 8-point algorithm.  Each image's 8 points are shifted to their centroid and scaled
// to mean distance sqrt(2); the 8 x 9 system [u x, u y, u, v x, v y, v, x, y, 1] f = 0 is
// reduced by Gaussian elimination with full pivoting (largest remaining |entry|, rows then
// columns), the one column left unpivoted is set to 1 and the rest back-substituted, so no
// entry of F is assumed non-zero; the normalized solution is projected to rank 2 (as cv2's
// models are: its 7-point solutions by construction, its 8-point one by SVD), then
// F = Tb^T F^ Ta.  Error: cv2's FM_RANSAC one, the larger of the two squared
// point-to-epipolar-line distances.
struct ModelF {
    static constexpr int K = 8;
    static constexpr bool refit = false;
    __device__ static bool solve(const double* xa, const double* ya, const double* xb, const double* yb, double* m) {
        double ca[2] = {0.0, 0.0}, cb[2] = {0.0, 0.0};
        for (int k = 0; k < 8; k++) {
            ca[0] += xa[k];
            ca[1] += ya[k];
            cb[0] += xb[k];
            cb[1] += yb[k];
        }
        for (int q = 0; q < 2; q++) {
            ca[q] /= 8.0;
            cb[q] /= 8.0;
        }
        double ra = 0.0, rb = 0.0;
        for (int k = 0; k < 8; k++) {
            ra += sqrt((xa[k] - ca[0]) * (xa[k] - ca[0]) + (ya[k] - ca[1]) * (ya[k] - ca[1]));
            rb += sqrt((xb[k] - cb[0]) * (xb[k] - cb[0]) + (yb[k] - cb[1]) * (yb[k] - cb[1]));
        }
        if (!(ra > 1e-12 && rb > 1e-12)) return false;
        const double sa = 1.4142135623730951 * 8.0 / ra, sb = 1.4142135623730951 * 8.0 / rb;
        double A[8][9];
        for (int k = 0; k < 8; k++) {
            const double x = (xa[k] - ca[0]) * sa, y = (ya[k] - ca[1]) * sa;
            const double u = (xb[k] - cb[0]) * sb, v = (yb[k] - cb[1]) * sb;
            double* r = A[k];
            r[0] = u * x; r[1] = u * y; r[2] = u;
            r[3] = v * x; r[4] = v * y; r[5] = v;
            r[6] = x; r[7] = y; r[8] = 1.0;
        }
        int perm[9];
#pragma unroll
        for (int k = 0; k < 9; k++) perm[k] = k;
        if (!f8_step<0>(A, perm) || !f8_step<1>(A, perm) || !f8_step<2>(A, perm) || !f8_step<3>(A, perm) ||
            !f8_step<4>(A, perm) || !f8_step<5>(A, perm) || !f8_step<6>(A, perm) || !f8_step<7>(A, perm))
            return false;
        double fz[9];   // solution in pivot order; the free (last) one is 1
        fz[8] = 1.0;
#pragma unroll
        for (int c = 7; c >= 0; c--) {
            double s = 0.0;
#pragma unroll
            for (int k = c + 1; k < 9; k++) s = s + A[c][k] * fz[k];
            fz[c] = -s / A[c][c];
        }
        double fh[9];   // fh[perm[k]] = fz[k]
#pragma unroll
        for (int j = 0; j < 9; j++) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < 9; k++) v = perm[k] == j ? fz[k] : v;
            fh[j] = v;
        }
        rank2_project(fh);
        // F = Tb^T F^ Ta with T = [[s, 0, -s cx], [0, s, -s cy], [0, 0, 1]]
        const double Ta[3][3] = {{sa, 0.0, -sa * ca[0]}, {0.0, sa, -sa * ca[1]}, {0.0, 0.0, 1.0}};
        const double Tb[3][3] = {{sb, 0.0, -sb * cb[0]}, {0.0, sb, -sb * cb[1]}, {0.0, 0.0, 1.0}};
        double G[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                G[i][j] = (fh[3 * i + 0] * Ta[0][j] + fh[3 * i + 1] * Ta[1][j]) + fh[3 * i + 2] * Ta[2][j];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) m[3 * i + j] = (Tb[0][i] * G[0][j] + Tb[1][i] * G[1][j]) + Tb[2][i] * G[2][j];
        return true;
    }
    // lines l2 = F x_a = (a2, b2, c2) in image b and l1 = F^T x_b = (a1, b1, .) in image a; the
    // residual d = x_b^T F x_a is the same number for both lines (cv2 evaluates it twice, as two
    // dot products equal in exact arithmetic), so it is formed once;
    // max(d^2 / n2, d^2 / n1) < thr2  <=>  d^2 < thr2 min(n1, n2) (n = |line normal|^2, a zero
    // normal never passes, as cv2's infinite error), fused multiply-adds in a fixed nesting, no
    // branch.  fmin returns the other operand for one NaN, as numpy's fmin.
    __device__ static bool inlier(const double* F, double x, double y, double u, double v, double thr2) {
        const double a2 = fma(F[0], x, fma(F[1], y, F[2])), b2 = fma(F[3], x, fma(F[4], y, F[5])),
                     c2 = fma(F[6], x, fma(F[7], y, F[8]));
        const double a1 = fma(F[0], u, fma(F[3], v, F[6])), b1 = fma(F[1], u, fma(F[4], v, F[7]));
        const double n2 = fma(a2, a2, b2 * b2), n1 = fma(a1, a1, b1 * b1);
        const double d = fma(u, a2, fma(v, b2, c2));
        return d * d < thr2 * fmin(n1, n2);
    }
};

// Similarity (the role of cv2.estimateAffinePartial2D in _similarity, matching.py:194-202):
// from 2 points, M = [[a, -b, tx], [b, a, ty]]; stored as m[0..5] = a, -b, tx, b, a, ty.
struct ModelS {
    static constexpr int K = 2;
    static constexpr bool refit = false;
    __device__ static bool solve(const double* xa, const double* ya, const double* xb, const double* yb, double* m) {
        const double dx = xa[1] - xa[0], dy = ya[1] - ya[0];
        const double ex = xb[1] - xb[0], ey = yb[1] - yb[0];
        const double den = dx * dx + dy * dy;
        if (!(den > 1e-12)) return false;
        const double a = (dx * ex + dy * ey) / den, b = (dx * ey - dy * ex) / den;
        m[0] = a;
        m[1] = -b;
        m[2] = xb[0] - (a * xa[0] - b * ya[0]);
        m[3] = b;
        m[4] = a;
        m[5] = yb[0] - (b * xa[0] + a * ya[0]);
        m[6] = 0.0;
        m[7] = 0.0;
        m[8] = 1.0;
        return true;
    }
    __device__ static bool inlier(const double* m, double x, double y, double u, double v, double thr2) {
        const double dx = fma(m[0], x, fma(m[1], y, m[2])) - u, dy = fma(m[3], x, fma(m[4], y, m[5])) - v;
        return fma(dx, dx, dy * dy) < thr2;
    }
};

// draw the K distinct matches of hypothesis hy of pair p (16 attempts per slot)
template <int K>
__device__ bool draw_sample(unsigned long long seed, int p, int hy, int n, int* idx) {
    for (int j = 0; j < K; j++) {
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
        if (cand < 0) return false;
        idx[j] = cand;
    }
    return true;
}

template <class M>
__device__ bool hypothesis(const double* A, const double* Bq, unsigned long long seed, int p, int hy, int n, double* m) {
    int idx[M::K];
    if (!draw_sample<M::K>(seed, p, hy, n, idx)) return false;
    double xa[M::K], ya[M::K], xb[M::K], yb[M::K];
    for (int j = 0; j < M::K; j++) {
        xa[j] = A[2 * idx[j]];
        ya[j] = A[2 * idx[j] + 1];
        xb[j] = Bq[2 * idx[j]];
        yb[j] = Bq[2 * idx[j] + 1];
    }
    return M::solve(xa, ya, xb, yb, m);
}

// seeds / thrs (device, per pair; may be null): pair p draws with (seeds[p], 0, h) instead of
// (seed, p, h) and tests against thrs[p]^2 instead of thr2, so its result depends on its own
// seed only, not on its position in the batch.
struct RansacPair {
    int p;                       // pair index of the draws
    unsigned long long seed;
    double thr2;
    long long o;                 // first match row
    int n;                       // matches
};

__device__ __forceinline__ RansacPair ransac_pair(int slot, const long long* __restrict__ off, double thr2_all,
                                                  unsigned long long seed_all, const unsigned long long* __restrict__ seeds,
                                                  const double* __restrict__ thrs) {
    RansacPair r;
    r.p = seeds ? 0 : slot;
    r.seed = seeds ? seeds[slot] : seed_all;
    r.thr2 = thrs ? thrs[slot] * thrs[slot] : thr2_all;
    r.o = off[slot];
    r.n = (int)(off[slot + 1] - r.o);
    return r;
}

// Scoring.  Key of hypothesis hy with model h: (inliers + 1) << 32 | (0xffffffff - hy), so the
// largest key has the most inliers, ties the lowest hypothesis; 0 when it gave no model.  Every
// lane tests its own hypothesis against the matches in order.  A match's index is the same for
// all lanes, so its four coordinates are read by scalar loads into SGPRs (no LDS staging, no
// barrier, no vector memory traffic) and used as the scalar operand of the float64
// instructions; four matches per iteration, one counter each (an add with carry-in per test).
template <class M>
__device__ __forceinline__ unsigned long long score_key(const RansacPair& R, const double* __restrict__ A,
                                                        const double* __restrict__ Bq, const double (&h)[9], bool ok,
                                                        int hy) {
    int c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    int k = 0;
    for (; k + 4 <= R.n; k += 4) {
        c0 += M::inlier(h, A[2 * k], A[2 * k + 1], Bq[2 * k], Bq[2 * k + 1], R.thr2);
        c1 += M::inlier(h, A[2 * k + 2], A[2 * k + 3], Bq[2 * k + 2], Bq[2 * k + 3], R.thr2);
        c2 += M::inlier(h, A[2 * k + 4], A[2 * k + 5], Bq[2 * k + 4], Bq[2 * k + 5], R.thr2);
        c3 += M::inlier(h, A[2 * k + 6], A[2 * k + 7], Bq[2 * k + 6], Bq[2 * k + 7], R.thr2);
    }
    for (; k < R.n; k++) c0 += M::inlier(h, A[2 * k], A[2 * k + 1], Bq[2 * k], Bq[2 * k + 1], R.thr2);
    const int cnt = (c0 + c1) + (c2 + c3);
    return ok ? ((unsigned long long)(cnt + 1) << 32) | (0xffffffffu - (unsigned)hy) : 0ull;
}

// the largest key of the block (256 threads)
__device__ unsigned long long block_max_key(unsigned long long key) {
    __shared__ unsigned long long s_key[256];
    const int tid = threadIdx.x;
    s_key[tid] = key;
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (tid < w) s_key[tid] = s_key[tid + w] > s_key[tid] ? s_key[tid + w] : s_key[tid];
        __syncthreads();
    }
    const unsigned long long best = s_key[0];
    __syncthreads();
    return best;
}

// the winner of key: rebuild its model (the same draws and solve), write the model, its inlier
// count and the inlier flags of every match (256 threads)
template <class M>
__device__ void ransac_finish(const RansacPair& R, int slot, const double* A, const double* Bq, unsigned long long key,
                              uint8_t* __restrict__ inlier, double* __restrict__ H_out, int* __restrict__ n_in) {
    __shared__ double s_h[9];
    __shared__ int s_have;
    const int tid = threadIdx.x;
    if (tid == 0) {
        const int bc = (int)(key >> 32) - 1, bh = (int)(0xffffffffu - (unsigned)key);
        double h[9];
        const bool have = bc >= 0 && hypothesis<M>(A, Bq, R.seed, R.p, bh, R.n, h);
        for (int k = 0; k < 9; k++) s_h[k] = have ? h[k] : 0.0;
        s_have = have;
        n_in[slot] = have ? bc : 0;
        for (int k = 0; k < 9; k++) H_out[9 * slot + k] = s_h[k];
    }
    __syncthreads();
    const bool have = s_have != 0;
    double h[9];
    for (int k = 0; k < 9; k++) h[k] = s_h[k];
    for (int i = tid; i < R.n; i += 256)
        inlier[R.o + i] = (uint8_t)(have && M::inlier(h, A[2 * i], A[2 * i + 1], Bq[2 * i], Bq[2 * i + 1], R.thr2));
}

// Scoring: block (p, r) scores hypotheses 256 r .. 256 r + 255 of pair p and merges its best
// key into word 0 of the pair's 9-double model slot (an atomic maximum; k_ransac_keys_zero
// clears the words first), so a pair with many matches spreads over as many blocks as it has
// rounds of 256 hypotheses.  k_ransac_pick reads the word back before it writes the model there.
__global__ void k_ransac_keys_zero(int n_pair, double* __restrict__ H_out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < n_pair) ((unsigned long long*)H_out)[9 * (long long)p] = 0ull;
}

// r0: the first round of the launch (hypotheses 256 (r0 + blockIdx.y) + threadIdx.x); live
// (optional, per pair): blocks of pairs whose flag is 0 exit at once (the adaptive form).
template <class M>
__global__ __launch_bounds__(256) void k_ransac_split(const long long* __restrict__ off, const double* __restrict__ pa,
                                                      const double* __restrict__ pb, int n_hyp, double thr2_all,
                                                      unsigned long long seed_all,
                                                      const unsigned long long* __restrict__ seeds,
                                                      const double* __restrict__ thrs, double* __restrict__ H_out,
                                                      int r0, const int* __restrict__ live) {
    const int slot = blockIdx.x;
    if (live && !live[slot]) return;
    const RansacPair R = ransac_pair(slot, off, thr2_all, seed_all, seeds, thrs);
    const double* A = pa + 2 * R.o;
    const double* Bq = pb + 2 * R.o;
    const int hy = 256 * (r0 + (int)blockIdx.y) + (int)threadIdx.x;
    double h[9];
#pragma unroll
    for (int k = 0; k < 9; k++) h[k] = 0.0;
    const bool ok = hy < n_hyp && R.n >= M::K && hypothesis<M>(A, Bq, R.seed, R.p, hy, R.n, h);
    const unsigned long long key = block_max_key(score_key<M>(R, A, Bq, h, ok, hy));
    if (threadIdx.x == 0 && key) atomicMax((unsigned long long*)H_out + 9 * (long long)slot, key);
}

// The same scoring in two launches, for the models whose solve needs many registers (the
// 4-point homography, the 8-point F: a fused kernel runs at the solve's occupancy).
// k_ransac_solve: block (x, r) solves hypotheses hy = 256 (r0 + r) .. + 255 of pair p0 + x into
// the plane layout hyp[q * plane + x * hstride + (hy - 256 r0)] (q = 0..8; valid[...] = 1 when
// the sample gave a model; hstride = the hypotheses one launch holds per pair);
// k_ransac_score's block (x, r) scores the same hypotheses.  live as in k_ransac_split.
template <class M>
__global__ __launch_bounds__(256) void k_ransac_solve(const long long* __restrict__ off, const double* __restrict__ pa,
                                                      const double* __restrict__ pb, int n_hyp,
                                                      unsigned long long seed_all,
                                                      const unsigned long long* __restrict__ seeds, int p0,
                                                      long long plane, double* __restrict__ hyp,
                                                      uint8_t* __restrict__ valid, int r0, int hstride,
                                                      const int* __restrict__ live) {
    const int hl = 256 * (int)blockIdx.y + (int)threadIdx.x;
    const int hy = 256 * r0 + hl;
    if (hy >= n_hyp || hl >= hstride) return;
    if (live && !live[p0 + (int)blockIdx.x]) return;
    const RansacPair R = ransac_pair(p0 + (int)blockIdx.x, off, 0.0, seed_all, seeds, nullptr);
    double h[9];
    const bool ok = R.n >= M::K && hypothesis<M>(pa + 2 * R.o, pb + 2 * R.o, R.seed, R.p, hy, R.n, h);
    const long long i = (long long)blockIdx.x * hstride + hl;
#pragma unroll
    for (int q = 0; q < 9; q++) hyp[q * plane + i] = ok ? h[q] : 0.0;
    valid[i] = ok ? 1 : 0;
}

template <class M>
__global__ __launch_bounds__(256) void k_ransac_score(const long long* __restrict__ off, const double* __restrict__ pa,
                                                      const double* __restrict__ pb, int n_hyp, double thr2_all,
                                                      unsigned long long seed_all,
                                                      const unsigned long long* __restrict__ seeds,
                                                      const double* __restrict__ thrs, int p0, long long plane,
                                                      const double* __restrict__ hyp,
                                                      const uint8_t* __restrict__ valid, double* __restrict__ H_out,
                                                      int r0, int hstride, const int* __restrict__ live) {
    const int slot = p0 + (int)blockIdx.x;
    if (live && !live[slot]) return;
    const RansacPair R = ransac_pair(slot, off, thr2_all, seed_all, seeds, thrs);
    const int hl = 256 * (int)blockIdx.y + (int)threadIdx.x;
    const int hy = 256 * r0 + hl;
    const long long i = (long long)blockIdx.x * hstride + hl;
    double h[9];
    bool ok = false;
#pragma unroll
    for (int q = 0; q < 9; q++) h[q] = 0.0;
    if (hy < n_hyp && hl < hstride) {
        ok = valid[i] != 0;
#pragma unroll
        for (int q = 0; q < 9; q++) h[q] = hyp[q * plane + i];
    }
    const unsigned long long key = block_max_key(score_key<M>(R, pa + 2 * R.o, pb + 2 * R.o, h, ok, hy));
    if (threadIdx.x == 0 && key) atomicMax((unsigned long long*)H_out + 9 * (long long)slot, key);
}

template <class M>
__global__ __launch_bounds__(256) void k_ransac_pick(const long long* __restrict__ off, const double* __restrict__ pa,
                                                     const double* __restrict__ pb, double thr2_all,
                                                     unsigned long long seed_all,
                                                     const unsigned long long* __restrict__ seeds,
                                                     const double* __restrict__ thrs, uint8_t* __restrict__ inlier,
                                                     double* __restrict__ H_out, int* __restrict__ n_in) {
    __shared__ unsigned long long s_best;
    const int slot = blockIdx.x;
    const RansacPair R = ransac_pair(slot, off, thr2_all, seed_all, seeds, thrs);
    if (threadIdx.x == 0) s_best = ((const unsigned long long*)H_out)[9 * (long long)slot];
    __syncthreads();                 // the key is read before ransac_finish overwrites the slot
    ransac_finish<M>(R, slot, pa + 2 * R.o, pb + 2 * R.o, s_best, inlier, H_out, n_in);
}

// Adaptive stopping, the rule by which cv2's RANSAC estimators end early: stop once the
// hypotheses drawn reach log(1 - confidence) / log(1 - w^K), w the inlier ratio of the best
// model so far, K its sample size.  Evaluated after each round of 256 hypotheses (R rounds
// done) as (1 - w^K)^(256 R) <= 1 - confidence, by multiplications only: w = c / n, w^K by
// K - 1 successive products, the 256th power by 8 squarings, then R - 1 products; the power
// underflowing to 0 means stop.  A restatement with the same operations decides identically.
template <int K>
__device__ __forceinline__ bool ransac_confident(int c, int n, int rounds, double omc) {
    const double w = (double)c / (double)n;
    double wk = w;
#pragma unroll
    for (int i = 1; i < K; i++) wk = wk * w;
    double q = 1.0 - wk;
#pragma unroll
    for (int i = 0; i < 8; i++) q = q * q;
    double y = q;
    for (int i = 1; i < rounds; i++) y = y * q;
    return y <= omc;
}

__global__ void k_ransac_adapt_init(int n_pair, double* __restrict__ H_out, int* __restrict__ live) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p < n_pair) {
        ((unsigned long long*)H_out)[9 * (long long)p] = 0ull;
        live[p] = 1;
    }
}

// After a round: done = hypotheses drawn per live pair so far (min(256 (r + 1), max_hyp)).  A
// live pair records done in used[] and stops when done reaches max_hyp, when it has fewer than
// K matches, or when the inlier count of its best key so far satisfies the rule above.
template <class M>
__global__ void k_ransac_adapt(int n_pair, const long long* __restrict__ off, int done, int max_hyp, double omc,
                               const double* __restrict__ H_out, int* __restrict__ live, int* __restrict__ used) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= n_pair || !live[p]) return;
    used[p] = done;
    const int n = (int)(off[p + 1] - off[p]);
    const unsigned long long key = ((const unsigned long long*)H_out)[9 * (long long)p];
    const int c = key ? (int)(key >> 32) - 1 : 0;
    if (done >= max_hyp || n < M::K || (c > 0 && ransac_confident<M::K>(c, n, done / 256, omc))) live[p] = 0;
}

// Least-squares refit of each pair's homography on the inliers of its winning hypothesis (the
// role of the Levenberg-Marquardt refinement cv2.findHomography runs on the RANSAC inliers):
// each image's inliers shifted to their centroid and scaled to mean distance sqrt(2), the
// normal equations of the DLT rows (h33 = 1) summed over the inliers, solved by Gaussian
// elimination with partial pivoting, then H = Tb^-1 H^ Ta scaled to h33 = 1.  The inlier flags
// and count stay those of the winning hypothesis.  A pair with fewer than 5 inliers, a pivot
// below 1e-12 of the largest diagonal entry, |h33| below 1e-12 of the largest entry or a
// non-finite result keeps the hypothesis' model.  Sums: lane l adds matches l, l + 64, ... in
// order (a non-inlier adds 0.0), then the 64 lane sums are added in lane order, so a CPU
// restatement gives the same bits.  One 64-lane block per pair.
constexpr int RF_LANES = 64;
constexpr int RF_TERMS = 44;     // the 36 upper-triangle entries of the 8 x 8 matrix, then the 8 of the right side

template <int M>
__device__ __forceinline__ void lane_sums(const double (&v)[M], double (*part)[RF_LANES + 1], double* tot) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int e = 0; e < M; e++) part[e][lane] = v[e];
    __syncthreads();
    for (int e = lane; e < M; e += RF_LANES) {
        double t = 0.0;
        for (int l = 0; l < RF_LANES; l++) t = t + part[e][l];
        tot[e] = t;
    }
    __syncthreads();
}

__global__ __launch_bounds__(RF_LANES) void k_refit_h(const long long* __restrict__ off, const double* __restrict__ pa,
                                                     const double* __restrict__ pb, const uint8_t* __restrict__ inlier,
                                                     const int* __restrict__ n_in, double* __restrict__ H_out) {
    __shared__ double s_part[RF_TERMS][RF_LANES + 1];
    __shared__ double s_tot[RF_TERMS];
    const int lane = threadIdx.x, p = blockIdx.x;
    const int cnt = n_in[p];
    if (cnt < 5) return;
    const long long o = off[p];
    const int n = (int)(off[p + 1] - o);
    const double* A = pa + 2 * o;
    const double* B = pb + 2 * o;
    const uint8_t* fl = inlier + o;
    const double dn = (double)cnt;
    double c4[4] = {0.0, 0.0, 0.0, 0.0};
    for (int i = lane; i < n; i += RF_LANES) {
        const bool in = fl[i] != 0;
        c4[0] = c4[0] + (in ? A[2 * i] : 0.0);
        c4[1] = c4[1] + (in ? A[2 * i + 1] : 0.0);
        c4[2] = c4[2] + (in ? B[2 * i] : 0.0);
        c4[3] = c4[3] + (in ? B[2 * i + 1] : 0.0);
    }
    lane_sums<4>(c4, s_part, s_tot);
    const double cax = s_tot[0] / dn, cay = s_tot[1] / dn, cbx = s_tot[2] / dn, cby = s_tot[3] / dn;
    double r2[2] = {0.0, 0.0};
    for (int i = lane; i < n; i += RF_LANES) {
        const bool in = fl[i] != 0;
        const double dxa = A[2 * i] - cax, dya = A[2 * i + 1] - cay;
        const double dxb = B[2 * i] - cbx, dyb = B[2 * i + 1] - cby;
        r2[0] = r2[0] + (in ? sqrt(dxa * dxa + dya * dya) : 0.0);
        r2[1] = r2[1] + (in ? sqrt(dxb * dxb + dyb * dyb) : 0.0);
    }
    lane_sums<2>(r2, s_part, s_tot);
    const double ra = s_tot[0], rb = s_tot[1];
    if (!(ra > 1e-12 && rb > 1e-12)) return;
    const double sa = 1.4142135623730951 * dn / ra, sb = 1.4142135623730951 * dn / rb;
    double acc[RF_TERMS];
#pragma unroll
    for (int e = 0; e < RF_TERMS; e++) acc[e] = 0.0;
    for (int i = lane; i < n; i += RF_LANES) {
        const bool in = fl[i] != 0;
        const double x = (A[2 * i] - cax) * sa, y = (A[2 * i + 1] - cay) * sa;
        const double u = (B[2 * i] - cbx) * sb, v = (B[2 * i + 1] - cby) * sb;
        const double r0[8] = {x, y, 1.0, 0.0, 0.0, 0.0, -(u * x), -(u * y)};
        const double r1[8] = {0.0, 0.0, 0.0, x, y, 1.0, -(v * x), -(v * y)};
        int e = 0;
#pragma unroll
        for (int a = 0; a < 8; a++)
#pragma unroll
            for (int b = a; b < 8; b++, e++) acc[e] = acc[e] + (in ? r0[a] * r0[b] + r1[a] * r1[b] : 0.0);
#pragma unroll
        for (int a = 0; a < 8; a++) acc[36 + a] = acc[36 + a] + (in ? r0[a] * u + r1[a] * v : 0.0);
    }
    lane_sums<RF_TERMS>(acc, s_part, s_tot);
    if (lane != 0) return;
    double N[8][9];
    {
        int e = 0;
        for (int a = 0; a < 8; a++)
            for (int b = a; b < 8; b++, e++) N[a][b] = N[b][a] = s_tot[e];
        for (int a = 0; a < 8; a++) N[a][8] = s_tot[36 + a];
    }
    double dmax = 0.0;
    for (int a = 0; a < 8; a++) dmax = fmax(dmax, N[a][a]);
    const double tol = 1e-12 * dmax;
    for (int c = 0; c < 8; c++) {
        int piv = c;
        double best = fabs(N[c][c]);
        for (int r = c + 1; r < 8; r++) {
            const double a = fabs(N[r][c]);
            if (a > best) {
                best = a;
                piv = r;
            }
        }
        if (!(best > tol)) return;
        if (piv != c)
            for (int k = 0; k < 9; k++) {
                const double t = N[c][k];
                N[c][k] = N[piv][k];
                N[piv][k] = t;
            }
        for (int r = c + 1; r < 8; r++) {
            const double f = N[r][c] / N[c][c];
            for (int k = c; k < 9; k++) N[r][k] = N[r][k] - f * N[c][k];
        }
    }
    double h[9];
    for (int c = 7; c >= 0; c--) {
        double s = N[c][8];
        for (int k = c + 1; k < 8; k++) s = s - N[c][k] * h[k];
        h[c] = s / N[c][c];
    }
    h[8] = 1.0;
    // H = Tb^-1 H^ Ta, Ta = [[sa, 0, -sa cax], [0, sa, -sa cay], [0, 0, 1]], Tb^-1 = [[1/sb, 0, cbx], [0, 1/sb, cby], [0, 0, 1]]
    const double Ta[3][3] = {{sa, 0.0, -sa * cax}, {0.0, sa, -sa * cay}, {0.0, 0.0, 1.0}};
    const double ib = 1.0 / sb;
    const double Tbi[3][3] = {{ib, 0.0, cbx}, {0.0, ib, cby}, {0.0, 0.0, 1.0}};
    double G[3][3], Hr[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) G[i][j] = (h[3 * i] * Ta[0][j] + h[3 * i + 1] * Ta[1][j]) + h[3 * i + 2] * Ta[2][j];
    double hmax = 0.0;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            Hr[3 * i + j] = (Tbi[i][0] * G[0][j] + Tbi[i][1] * G[1][j]) + Tbi[i][2] * G[2][j];
            hmax = fmax(hmax, fabs(Hr[3 * i + j]));
        }
    const double h33 = Hr[8];
    if (!(fabs(h33) > 1e-12 * hmax)) return;
    bool finite = true;
    for (int k = 0; k < 9; k++) {
        Hr[k] = Hr[k] / h33;
        finite &= isfinite(Hr[k]);
    }
    if (!finite) return;
    for (int k = 0; k < 9; k++) H_out[9 * p + k] = Hr[k];
}

// out[k] = xy[idx[k]] (float64 (x, y) rows) for k < n; an index outside [0, n_xy) gives NaN.
__global__ __launch_bounds__(256) void k_gather_xy(long long n, const int* __restrict__ idx,
                                                   const double* __restrict__ xy, long long n_xy,
                                                   double* __restrict__ out) {
    const long long k = (long long)blockIdx.x * 256 + threadIdx.x;
    if (k >= n) return;
    const int i = idx[k];
    const bool ok = i >= 0 && (long long)i < n_xy;
    const double2 v = ok ? ((const double2*)xy)[i] : make_double2(__builtin_nan(""), __builtin_nan(""));
    ((double2*)out)[k] = v;
}

// Lists of matches moved between the chain's stages on the device (one 256-thread block per
// list): list L's rows src_off[L] .. + cnt[L] of q_src / t_src (those with keep[row] != 0 when
// keep is given, in order) go to rows dst_off[L] .. of q_dst / t_dst, and, when pa_dst / pb_dst
// are given, the positions xy[base_a[L] + q] / xy[base_b[L] + t] to the same rows (NaN for a row
// outside [0, n_xy)); rows at or past dst_cap are not written; kept[L] = rows taken.
__global__ __launch_bounds__(256) void k_gather_lists(const long long* __restrict__ src_off,
                                                      const int* __restrict__ cnt,
                                                      const long long* __restrict__ dst_off, long long dst_cap,
                                                      const long long* __restrict__ base_a,
                                                      const long long* __restrict__ base_b,
                                                      const int* __restrict__ q_src, const int* __restrict__ t_src,
                                                      const uint8_t* __restrict__ keep,
                                                      const double* __restrict__ xy, long long n_xy,
                                                      int* __restrict__ q_dst, int* __restrict__ t_dst,
                                                      double* __restrict__ pa_dst, double* __restrict__ pb_dst,
                                                      int* __restrict__ kept) {
    __shared__ int s_w[4];
    const int L = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const long long so = src_off[L], d0 = dst_off[L];
    const int n = cnt[L];
    const long long ba = base_a ? base_a[L] : 0, bb = base_b ? base_b[L] : 0;
    const double2 nan2 = make_double2(__builtin_nan(""), __builtin_nan(""));
    int carry = 0;
    for (int b0 = 0; b0 < n; b0 += 256) {
        const int i = b0 + tid;
        const bool in = i < n && (!keep || keep[so + i] != 0);
        const unsigned long long bal = __ballot(in);
        if (lane == 0) s_w[wv] = __popcll(bal);
        __syncthreads();
        int pre = 0, tot = 0;
#pragma unroll
        for (int w2 = 0; w2 < 4; w2++) {
            pre += w2 < wv ? s_w[w2] : 0;
            tot += s_w[w2];
        }
        __syncthreads();                 // s_w is read before the next round writes it
        const long long d = d0 + carry + pre + __popcll(bal & ((1ull << lane) - 1ull));
        if (in && d >= 0 && d < dst_cap) {
            const int q = q_src[so + i], t = t_src[so + i];
            if (q_dst) q_dst[d] = q;
            if (t_dst) t_dst[d] = t;
            if (pa_dst) {
                const long long r = ba + q;
                ((double2*)pa_dst)[d] = r >= 0 && r < n_xy ? ((const double2*)xy)[r] : nan2;
            }
            if (pb_dst) {
                const long long r = bb + t;
                ((double2*)pb_dst)[d] = r >= 0 && r < n_xy ? ((const double2*)xy)[r] : nan2;
            }
        }
        carry += tot;
    }
    if (kept && tid == 0) kept[L] = carry;
}

// Least-squares similarity b ~ s R(theta) a + t of each pair's inlier rows (one 256-thread block
// per pair): centroids, then over the centred rows den = sum |a|^2, c = sum a.b / den,
// s = sum a x b / den, t = mb - [[c, -s], [s, c]] ma; out[4p..4p+3] = (atan2(s, c), hypot(c, s),
// tx, ty), or NaN when fewer than 2 inliers or den <= 1e-12 n (the points coincide).  Sums:
// per thread over rows tid, tid + 256, ... in order, then a fixed tree over the threads.
__device__ __forceinline__ void block_sum5(double* v, double (*red)[257]) {
    const int tid = threadIdx.x;
#pragma unroll
    for (int e = 0; e < 5; e++) red[e][tid] = v[e];
    __syncthreads();
    for (int w = 128; w >= 1; w >>= 1) {
        if (tid < w)
#pragma unroll
            for (int e = 0; e < 5; e++) red[e][tid] = red[e][tid] + red[e][tid + w];
        __syncthreads();
    }
#pragma unroll
    for (int e = 0; e < 5; e++) v[e] = red[e][0];
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_similarity_fit(const long long* __restrict__ off,
                                                        const double* __restrict__ pa, const double* __restrict__ pb,
                                                        const uint8_t* __restrict__ inlier, double* __restrict__ out) {
    __shared__ double red[5][257];
    const int p = blockIdx.x, tid = threadIdx.x;
    const long long o = off[p];
    const int n = (int)(off[p + 1] - o);
    const double* A = pa + 2 * o;
    const double* B = pb + 2 * o;
    const uint8_t* f = inlier + o;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = tid; i < n; i += 256)
        if (f[i]) {
            v[0] = v[0] + A[2 * i];
            v[1] = v[1] + A[2 * i + 1];
            v[2] = v[2] + B[2 * i];
            v[3] = v[3] + B[2 * i + 1];
            v[4] = v[4] + 1.0;
        }
    block_sum5(v, red);
    const double cnt = v[4];
    const double nan = __builtin_nan("");
    if (!(cnt >= 2.0)) {
        if (tid < 4) out[4 * p + tid] = nan;
        return;
    }
    const double max_ = v[0] / cnt, may = v[1] / cnt, mbx = v[2] / cnt, mby = v[3] / cnt;
    double w[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int i = tid; i < n; i += 256)
        if (f[i]) {
            const double ax = A[2 * i] - max_, ay = A[2 * i + 1] - may;
            const double bx = B[2 * i] - mbx, by = B[2 * i + 1] - mby;
            w[0] = w[0] + (ax * ax + ay * ay);
            w[1] = w[1] + (ax * bx + ay * by);
            w[2] = w[2] + (ax * by - ay * bx);
        }
    block_sum5(w, red);
    if (tid != 0) return;
    const double den = w[0];
    if (!(den > 1e-12 * cnt)) {
        for (int e = 0; e < 4; e++) out[4 * p + e] = nan;
        return;
    }
    const double c = w[1] / den, sn = w[2] / den;
    out[4 * p] = atan2(sn, c);
    out[4 * p + 1] = hypot(c, sn);
    out[4 * p + 2] = mbx - (c * max_ - sn * may);
    out[4 * p + 3] = mby - (sn * max_ + c * may);
}

}  // namespace fsx

using namespace fsx;

template <class M>
static void launch_refit(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                         const uint8_t* inlier_dev, const int* n_in_dev, double* M_dev, hipStream_t stream) {
    if (M::refit)
        hipLaunchKernelGGL(k_refit_h, dim3(n_pair), dim3(RF_LANES), 0, stream, off_dev, pa_dev, pb_dev, inlier_dev,
                           n_in_dev, M_dev);
}

// Hypothesis scoring and the winner: the keys cleared, one block per pair and round of 256
// hypotheses, then one block per pair rebuilds the winner (most inliers, then the lowest
// hypothesis) and labels the matches.  The homography and F solve into a stream-ordered
// temporary first (pairs in chunks of at most RANSAC_HYP_BYTES of hypotheses), the similarity
// solves inside its scoring kernel.
constexpr long long RANSAC_HYP_BYTES = 256ll << 20;

template <class M>
static int launch_score(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev, int n_hyp,
                        double thr2, unsigned long long seed, const unsigned long long* seed_dev, const double* thr_dev,
                        uint8_t* inlier_dev, double* M_dev, int* n_in_dev, hipStream_t stream) {
    const int rounds = (n_hyp + 255) / 256;
    hipLaunchKernelGGL(k_ransac_keys_zero, dim3((unsigned)((n_pair + 255) / 256)), dim3(256), 0, stream, n_pair, M_dev);
    if constexpr (M::K <= 2) {
        hipLaunchKernelGGL(k_ransac_split<M>, dim3(n_pair, rounds), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                           n_hyp, thr2, seed, seed_dev, thr_dev, M_dev, 0, (const int*)nullptr);
    } else {
        const long long per_pair = (long long)n_hyp * (9 * sizeof(double) + 1);
        const int chunk = (int)std::max(1ll, std::min((long long)n_pair, RANSAC_HYP_BYTES / per_pair));
        const long long plane = (long long)chunk * n_hyp;
        void* tmp = nullptr;
        FSX_HIP(hipMallocAsync(&tmp, (size_t)(plane * (9 * sizeof(double) + 1)), stream));
        double* hyp = (double*)tmp;
        uint8_t* valid = (uint8_t*)(hyp + 9 * plane);
        for (int p0 = 0; p0 < n_pair; p0 += chunk) {
            const int cp = std::min(chunk, n_pair - p0);
            hipLaunchKernelGGL(k_ransac_solve<M>, dim3(cp, rounds), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                               n_hyp, seed, seed_dev, p0, plane, hyp, valid, 0, n_hyp, (const int*)nullptr);
            hipLaunchKernelGGL(k_ransac_score<M>, dim3(cp, rounds), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                               n_hyp, thr2, seed, seed_dev, thr_dev, p0, plane, hyp, valid, M_dev, 0, n_hyp,
                               (const int*)nullptr);
        }
        FSX_HIP(hipFreeAsync(tmp, stream));
    }
    hipLaunchKernelGGL(k_ransac_pick<M>, dim3(n_pair), dim3(256), 0, stream, off_dev, pa_dev, pb_dev, thr2, seed,
                       seed_dev, thr_dev, inlier_dev, M_dev, n_in_dev);
    return FSX_OK;
}

template <class M>
static int launch_ransac(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev, int n_hyp,
                         double thr, unsigned long long seed, uint8_t* inlier_dev, double* H_dev, int* n_in_dev,
                         void* stream) {
    FSX_REQUIRE(n_pair >= 0 && n_hyp >= 1 && n_hyp <= (1 << 20) && thr > 0, "bad RANSAC arguments");
    if (n_pair == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && pa_dev && pb_dev && inlier_dev && H_dev && n_in_dev, "null pointer argument");
    const int rc = launch_score<M>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr * thr, seed, nullptr, nullptr,
                                   inlier_dev, H_dev, n_in_dev, (hipStream_t)stream);
    if (rc != FSX_OK) return rc;
    launch_refit<M>(n_pair, off_dev, pa_dev, pb_dev, inlier_dev, n_in_dev, H_dev, (hipStream_t)stream);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

template <class M>
static int launch_ransac_per_pair(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                                  int n_hyp, const double* thr_dev, const unsigned long long* seed_dev,
                                  uint8_t* inlier_dev, double* M_dev, int* n_in_dev, void* stream) {
    FSX_REQUIRE(n_pair >= 0 && n_hyp >= 1 && n_hyp <= (1 << 20), "bad RANSAC arguments");
    if (n_pair == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && pa_dev && pb_dev && thr_dev && seed_dev && inlier_dev && M_dev && n_in_dev,
                "null pointer argument");
    const int rc = launch_score<M>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, 0.0, 0ull, seed_dev, thr_dev, inlier_dev,
                                   M_dev, n_in_dev, (hipStream_t)stream);
    if (rc != FSX_OK) return rc;
    launch_refit<M>(n_pair, off_dev, pa_dev, pb_dev, inlier_dev, n_in_dev, M_dev, (hipStream_t)stream);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// The adaptive form for pairs with their own thresholds and seeds: rounds of 256 hypotheses,
// each launched for every pair but run only by the pairs still live, k_ransac_adapt after each
// round; the winner is the best key over the rounds a pair ran (most inliers, then the lowest
// hypothesis), picked, labelled and refit as in the fixed form.  One stream-ordered temporary
// holds the flags and one round's hypotheses (homography, F).
template <class M>
static int launch_ransac_adaptive(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                                  int max_hyp, double conf, const double* thr_dev,
                                  const unsigned long long* seed_dev, uint8_t* inlier_dev, double* M_dev,
                                  int* n_in_dev, int* used_dev, hipStream_t stream) {
    const int rounds = (max_hyp + 255) / 256;
    const double omc = 1.0 - conf;
    const unsigned pair_blocks = (unsigned)((n_pair + 255) / 256);
    const long long per_pair = 256ll * (9 * sizeof(double) + 1);
    const int chunk = M::K <= 2 ? 0 : (int)std::max(1ll, std::min((long long)n_pair, RANSAC_HYP_BYTES / per_pair));
    const long long plane = (long long)chunk * 256;
    const size_t flag_bytes = ((size_t)n_pair * 2 * sizeof(int) + 255) & ~(size_t)255;
    void* tmp = nullptr;
    FSX_HIP(hipMallocAsync(&tmp, flag_bytes + (size_t)(plane * (9 * sizeof(double) + 1)), stream));
    int* live = (int*)tmp;
    int* used = used_dev ? used_dev : live + n_pair;
    double* hyp = (double*)((char*)tmp + flag_bytes);
    uint8_t* valid = (uint8_t*)(hyp + 9 * plane);
    hipLaunchKernelGGL(k_ransac_adapt_init, dim3(pair_blocks), dim3(256), 0, stream, n_pair, M_dev, live);
    for (int r = 0; r < rounds; r++) {
        if constexpr (M::K <= 2) {
            hipLaunchKernelGGL(k_ransac_split<M>, dim3(n_pair, 1), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                               max_hyp, 0.0, 0ull, seed_dev, thr_dev, M_dev, r, (const int*)live);
        } else {
            for (int p0 = 0; p0 < n_pair; p0 += chunk) {
                const int cp = std::min(chunk, n_pair - p0);
                hipLaunchKernelGGL(k_ransac_solve<M>, dim3(cp, 1), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                                   max_hyp, 0ull, seed_dev, p0, plane, hyp, valid, r, 256, (const int*)live);
                hipLaunchKernelGGL(k_ransac_score<M>, dim3(cp, 1), dim3(256), 0, stream, off_dev, pa_dev, pb_dev,
                                   max_hyp, 0.0, 0ull, seed_dev, thr_dev, p0, plane, hyp, valid, M_dev, r, 256,
                                   (const int*)live);
            }
        }
        hipLaunchKernelGGL(k_ransac_adapt<M>, dim3(pair_blocks), dim3(256), 0, stream, n_pair, off_dev,
                           std::min(256 * (r + 1), max_hyp), max_hyp, omc, (const double*)M_dev, live, used);
    }
    hipLaunchKernelGGL(k_ransac_pick<M>, dim3(n_pair), dim3(256), 0, stream, off_dev, pa_dev, pb_dev, 0.0, 0ull,
                       seed_dev, thr_dev, inlier_dev, M_dev, n_in_dev);
    FSX_HIP(hipFreeAsync(tmp, stream));
    launch_refit<M>(n_pair, off_dev, pa_dev, pb_dev, inlier_dev, n_in_dev, M_dev, stream);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

extern "C" {

// Deterministic RANSAC for n_pair sets of matches: pair p's matches are rows off_dev[p] ..
// off_dev[p+1] of pa_dev / pb_dev (float64 (x, y) in image a / image b).  n_hyp hypotheses
// per pair (hypothesis h draws its sample with a counter hash of (seed, p, h)); the one with
// the most inliers wins (ties: lowest h).  Writes inlier_dev (one byte per match) and n_in_dev
// (per pair) from the winner, and H_dev (9 float64 per pair, all zero when no model): the
// winner refit on its inliers by k_refit_h.
int fsx_ransac_homography(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev, int n_hyp,
                          double thr, unsigned long long seed, uint8_t* inlier_dev, double* H_dev, int* n_in_dev,
                          void* stream) {
    return launch_ransac<ModelH>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr, seed, inlier_dev, H_dev, n_in_dev,
                                 stream);
}

int fsx_ransac_fundamental(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                           int n_hyp, double thr, unsigned long long seed, uint8_t* inlier_dev, double* F_dev,
                           int* n_in_dev, void* stream) {
    return launch_ransac<ModelF>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr, seed, inlier_dev, F_dev, n_in_dev,
                                 stream);
}

int fsx_ransac_similarity(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                          int n_hyp, double thr, unsigned long long seed, uint8_t* inlier_dev, double* M_dev,
                          int* n_in_dev, void* stream) {
    return launch_ransac<ModelS>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr, seed, inlier_dev, M_dev, n_in_dev,
                                 stream);
}


// One RANSAC launch for pairs that each carry their own threshold and seed (device arrays
// thr_dev, seed_dev of n_pair entries): model 0 = homography, 1 = fundamental matrix,
// 2 = similarity.  Pair p draws hypothesis h from the counter hash of (seed_dev[p], 0, h), so
// its result is the one a single-pair call with that seed gives, wherever p sits in the batch.
int fsx_ransac_batch(int model, int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                     int n_hyp, const double* thr_dev, const unsigned long long* seed_dev, uint8_t* inlier_dev,
                     double* M_dev, int* n_in_dev, void* stream) {
    switch (model) {
        case 0:
            return launch_ransac_per_pair<ModelH>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr_dev, seed_dev,
                                                  inlier_dev, M_dev, n_in_dev, stream);
        case 1:
            return launch_ransac_per_pair<ModelF>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr_dev, seed_dev,
                                                  inlier_dev, M_dev, n_in_dev, stream);
        case 2:
            return launch_ransac_per_pair<ModelS>(n_pair, off_dev, pa_dev, pb_dev, n_hyp, thr_dev, seed_dev,
                                                  inlier_dev, M_dev, n_in_dev, stream);
        default:
            FSX_REQUIRE(false, "model must be 0 (homography), 1 (fundamental) or 2 (similarity)");
    }
}

// fsx_ransac_batch with adaptive stopping: at most max_hyp hypotheses per pair, drawn in rounds
// of 256 until the inlier ratio w of the pair's best model so far gives
// (1 - w^K)^(hypotheses drawn) <= 1 - confidence (K = 4, 8, 2 for models 0, 1, 2).  used_dev
// (optional, int per pair): hypotheses drawn.  A pair's result depends on its own seed and
// matches only.
int fsx_ransac_batch_adaptive(int model, int n_pair, const long long* off_dev, const double* pa_dev,
                              const double* pb_dev, int max_hyp, double confidence, const double* thr_dev,
                              const unsigned long long* seed_dev, uint8_t* inlier_dev, double* M_dev, int* n_in_dev,
                              int* used_dev, void* stream) {
    FSX_REQUIRE(n_pair >= 0 && max_hyp >= 1 && max_hyp <= (1 << 20), "bad RANSAC arguments");
    FSX_REQUIRE(confidence > 0.0 && confidence < 1.0, "confidence must lie in (0, 1)");
    if (n_pair == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && pa_dev && pb_dev && thr_dev && seed_dev && inlier_dev && M_dev && n_in_dev,
                "null pointer argument");
    const hipStream_t st = (hipStream_t)stream;
    switch (model) {
        case 0:
            return launch_ransac_adaptive<ModelH>(n_pair, off_dev, pa_dev, pb_dev, max_hyp, confidence, thr_dev,
                                                  seed_dev, inlier_dev, M_dev, n_in_dev, used_dev, st);
        case 1:
            return launch_ransac_adaptive<ModelF>(n_pair, off_dev, pa_dev, pb_dev, max_hyp, confidence, thr_dev,
                                                  seed_dev, inlier_dev, M_dev, n_in_dev, used_dev, st);
        case 2:
            return launch_ransac_adaptive<ModelS>(n_pair, off_dev, pa_dev, pb_dev, max_hyp, confidence, thr_dev,
                                                  seed_dev, inlier_dev, M_dev, n_in_dev, used_dev, st);
        default:
            FSX_REQUIRE(false, "model must be 0 (homography), 1 (fundamental) or 2 (similarity)");
    }
}

// out_dev[k] (two float64) = xy_dev[idx_dev[k]] for k < n; an index outside [0, n_xy) gives
// NaN.  Gathers the matched positions of many pairs from all photos' positions in one launch.
int fsx_gather_xy(long long n, const int* idx_dev, const double* xy_dev, long long n_xy, double* out_dev,
                  void* stream) {
    FSX_REQUIRE(n >= 0 && n_xy >= 0, "bad gather sizes");
    if (n == 0) return FSX_OK;
    FSX_REQUIRE(idx_dev && xy_dev && out_dev, "null pointer argument");
    FSX_REQUIRE(((uintptr_t)xy_dev & 15) == 0 && ((uintptr_t)out_dev & 15) == 0, "xy / out must be 16-byte aligned");
    hipLaunchKernelGGL(k_gather_xy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream, n, idx_dev,
                       xy_dev, n_xy, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// The chain's lists of matches moved on the device between its stages: list L's cnt_dev[L]
// rows from src_off_dev[L] of q_src_dev / t_src_dev (int32 feature indices; with keep_dev, only
// rows whose byte is non-zero, in order) to rows dst_off_dev[L] .. of q_dst_dev / t_dst_dev, and
// optionally their positions xy_dev[base_a_dev[L] + q] / xy_dev[base_b_dev[L] + t] (float64
// (x, y), NaN outside [0, n_xy)) to pa_dst_dev / pb_dst_dev; rows at or past dst_cap are
// skipped; kept_dev[L] (optional) = rows taken.  Any of q_dst, t_dst, pa_dst, pb_dst may be
// null.  xy and the position outputs 16-byte aligned.
int fsx_gather_lists(int n_list, const long long* src_off_dev, const int* cnt_dev, const long long* dst_off_dev,
                     long long dst_cap, const long long* base_a_dev, const long long* base_b_dev,
                     const int* q_src_dev, const int* t_src_dev, const uint8_t* keep_dev, const double* xy_dev,
                     long long n_xy, int* q_dst_dev, int* t_dst_dev, double* pa_dst_dev, double* pb_dst_dev,
                     int* kept_dev, void* stream) {
    FSX_REQUIRE(n_list >= 0 && dst_cap >= 0 && n_xy >= 0, "bad list arguments");
    if (n_list == 0) return FSX_OK;
    FSX_REQUIRE(src_off_dev && cnt_dev && dst_off_dev && q_src_dev && t_src_dev, "null pointer argument");
    FSX_REQUIRE((!pa_dst_dev || base_a_dev) && (!pb_dst_dev || base_b_dev) && (!(pa_dst_dev || pb_dst_dev) || xy_dev),
                "positions need xy_dev and their row bases");
    FSX_REQUIRE(((uintptr_t)xy_dev & 15) == 0 && ((uintptr_t)pa_dst_dev & 15) == 0 && ((uintptr_t)pb_dst_dev & 15) == 0,
                "xy / positions must be 16-byte aligned");
    hipLaunchKernelGGL(k_gather_lists, dim3((unsigned)n_list), dim3(256), 0, (hipStream_t)stream, src_off_dev, cnt_dev,
                       dst_off_dev, dst_cap, base_a_dev, base_b_dev, q_src_dev, t_src_dev, keep_dev, xy_dev, n_xy,
                       q_dst_dev, t_dst_dev, pa_dst_dev, pb_dst_dev, kept_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Least-squares similarity (rotation, scale, translation) of each pair's inlier rows: pair p's
// rows are off_dev[p] .. off_dev[p+1] of pa_dev / pb_dev (float64 (x, y)), inlier_dev one byte
// per row; out_dev[4p .. 4p+3] = (angle rad, scale, tx, ty) of b ~ s R a + t, NaN when the pair
// has fewer than 2 inliers or they coincide.
int fsx_similarity_fit_batch(int n_pair, const long long* off_dev, const double* pa_dev, const double* pb_dev,
                             const uint8_t* inlier_dev, double* out_dev, void* stream) {
    FSX_REQUIRE(n_pair >= 0, "bad pair count");
    if (n_pair == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && pa_dev && pb_dev && inlier_dev && out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_similarity_fit, dim3(n_pair), dim3(256), 0, (hipStream_t)stream, off_dev, pa_dev, pb_dev,
                       inlier_dev, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
