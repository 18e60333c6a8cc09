// Keypoint geometry around the matcher on gfx950:
//   * feature positions to full-resolution undistorted pixels (matching.py:29-35
//     load_features' rescale and matching.py:48-64 _get's camera.undistort_pixels,
//     camera.py:5-29), bit-exact to the numpy expression order;
//   * the GNSS/gimbal ground-plane prior (matching.py:155-172 _predict);
//   * candidate pair selection from GPS positions (pipeline.py:452-476: k nearest within
//     2*half_long*0.95, plus temporal neighbours i+1, i+2).
// All arithmetic is float64 without contraction (the build passes -ffp-contract=off), in the
// operation order of the numpy expressions it restates.
#include "fsx_common.h"
#include <math.h>
#include <string.h>

namespace fsx {

// ------------------------------------------------------------------------------------------
// Undistortion

struct UdImage {
    const float* xy_in;     // (n, 2) float32 feature positions in reduced-image pixels
    double* xy_out;         // (n, 2) full-resolution pixels (may be null)
    double* und_out;        // (n, 2) undistorted pixels
    int n;
    int iters;              // fixed-point iterations (camera.py:5 default 25)
    int distorted;          // 0: all coefficients zero (camera.py:14-15 early return)
    int pad;
    double sx, sy;          // W / rw, H / rh (matching.py:33)
    double f, cx, cy, k1, k2, p1, p2, k3;
};

// camera.py:5-26 undistort_normalized then camera.py:28-30 back to pixels with the same f,
// cx, cy.
__device__ __forceinline__ void undistort_px(const UdImage& U, double px, double py, double& ux, double& uy) {
    const double xd = (px - U.cx) / U.f;
    const double yd = (py - U.cy) / U.f;
    double x = xd, y = yd;
    if (U.distorted) {
        for (int it = 0; it < U.iters; it++) {
            const double r2 = x * x + y * y;
            double rad = ((1.0 + U.k1 * r2) + (U.k2 * r2) * r2) + ((U.k3 * r2) * r2) * r2;
            rad = fabs(rad) < 0.2 ? 0.2 : rad;
            const double dx = ((2.0 * U.p1) * x) * y + U.p2 * (r2 + (2.0 * x) * x);
            const double dy = U.p1 * (r2 + (2.0 * y) * y) + ((2.0 * U.p2) * x) * y;
            x = (xd - dx) / rad;
            y = (yd - dy) / rad;
        }
    }
    ux = x * U.f + U.cx;
    uy = y * U.f + U.cy;
}

__global__ __launch_bounds__(256) void k_undistort(const UdImage* __restrict__ tab) {
    const UdImage& U = tab[blockIdx.y];
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= U.n) return;
    // matching.py:33: (xy + 0.5) * [W/rw, H/rh] - 0.5 in float64
    const double px = ((double)U.xy_in[2 * i] + 0.5) * U.sx - 0.5;
    const double py = ((double)U.xy_in[2 * i + 1] + 0.5) * U.sy - 0.5;
    if (U.xy_out) {
        U.xy_out[2 * i] = px;
        U.xy_out[2 * i + 1] = py;
    }
    undistort_px(U, px, py, U.und_out[2 * i], U.und_out[2 * i + 1]);
}

// float64 pixel input (camera.undistort_pixels called directly)
__global__ __launch_bounds__(256) void k_undistort_f64(int n, const double* __restrict__ uv, UdImage U,
                                                       double* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    undistort_px(U, uv[2 * i], uv[2 * i + 1], out[2 * i], out[2 * i + 1]);
}

static UdImage ud_params(const double* k, int iters) {
    UdImage U;
    memset(&U, 0, sizeof(U));
    U.iters = iters;
    U.sx = U.sy = 1.0;
    U.f = k[0];
    U.cx = k[1];
    U.cy = k[2];
    U.k1 = k[3];
    U.k2 = k[4];
    U.p1 = k[5];
    U.p2 = k[6];
    U.k3 = k[7];
    U.distorted = !(U.k1 == 0 && U.k2 == 0 && U.k3 == 0 && U.p1 == 0 && U.p2 == 0);
    return U;
}

// ------------------------------------------------------------------------------------------
// Ground-plane prior

struct GroundCams {
    double Kinv_i[9];   // inverse intrinsics of camera i (row-major)
    double R_i[9];      // world -> camera rotation of i
    double C_i[3];      // camera centre of i
    double zg_i;        // ground height under i
    double R_j[9];
    double K_j[9];
    double C_j[3];
};

static_assert(sizeof(GroundCams) == 43 * sizeof(double), "GroundCams is 43 packed doubles");

// matching.py:155-172 for one undistorted pixel (u0, u1) of camera i: returns ok (t > 0 and
// z > 0) and the predicted pixel in camera j
__device__ __forceinline__ bool ground_pred(const GroundCams& P, double u0, double u1, double& px, double& py) {
    // rays = [u, 1] @ inv(K_i)^T ; d = rays @ R_i
    double ray[3];
#pragma unroll
    for (int m = 0; m < 3; m++) ray[m] = (u0 * P.Kinv_i[3 * m] + u1 * P.Kinv_i[3 * m + 1]) + 1.0 * P.Kinv_i[3 * m + 2];
    double d[3];
#pragma unroll
    for (int c = 0; c < 3; c++) d[c] = (ray[0] * P.R_i[c] + ray[1] * P.R_i[3 + c]) + ray[2] * P.R_i[6 + c];
    double dz = d[2];
    dz = fabs(dz) < 1e-6 ? -1e-6 : dz;
    const double t = (P.zg_i - P.C_i[2]) / dz;
    double v[3];
#pragma unroll
    for (int c = 0; c < 3; c++) v[c] = (P.C_i[c] + t * d[c]) - P.C_j[c];
    // q = (G - C_j) @ R_j^T @ K_j^T
    double w[3];
#pragma unroll
    for (int a = 0; a < 3; a++) w[a] = (v[0] * P.R_j[3 * a] + v[1] * P.R_j[3 * a + 1]) + v[2] * P.R_j[3 * a + 2];
    double q[3];
#pragma unroll
    for (int b = 0; b < 3; b++) q[b] = (w[0] * P.K_j[3 * b] + w[1] * P.K_j[3 * b + 1]) + w[2] * P.K_j[3 * b + 2];
    double z = q[2];
    const bool ok = (t > 0.0) && (z > 0.0);
    z = fabs(z) < 1e-9 ? 1e-9 : z;
    px = q[0] / z;
    py = q[1] / z;
    return ok;
}

__global__ __launch_bounds__(256) void k_ground_predict(int n, const double* __restrict__ u, GroundCams P,
                                                        double* __restrict__ pred, uint8_t* __restrict__ ok) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    ok[i] = (uint8_t)ground_pred(P, u[2 * i], u[2 * i + 1], pred[2 * i], pred[2 * i + 1]);
}

// The prior filter of match_pair (matching.py:222-225, :238-241) for many pairs: row r of pair
// p (rows off[p] .. off[p+1]) keeps its match iff the ground-plane prediction of ua[r] through
// cams[p] is valid and lands within tol[p] of ub[r] (Euclidean norm, float64).  tol[p] < 0:
// the pair uses no prior and keeps every row.
__global__ __launch_bounds__(256) void k_prior_keep(const long long* __restrict__ off, const double* __restrict__ ua,
                                                    const double* __restrict__ ub, const GroundCams* __restrict__ cams,
                                                    const double* __restrict__ tol, uint8_t* __restrict__ keep) {
    const int p = blockIdx.y;
    const long long r = off[p] + (long long)blockIdx.x * 256 + threadIdx.x;
    if (r >= off[p + 1]) return;
    const double tp = tol[p];
    if (tp < 0.0) {
        keep[r] = 1;
        return;
    }
    double px, py;
    const bool ok = ground_pred(cams[p], ua[2 * r], ua[2 * r + 1], px, py);
    const double dx = px - ub[2 * r], dy = py - ub[2 * r + 1];
    keep[r] = (uint8_t)(ok && sqrt(dx * dx + dy * dy) < tp);
}

// ------------------------------------------------------------------------------------------
// Pair selection: one block per photo i.  The k+1 nearest photos (Euclidean GPS distance,
// ties to the lower index; the first is normally i itself and is skipped, as the reference's
// dists[i][1:]) are found by k+1 rounds of a block-wide lexicographic argmin over
// (distance bits, index) above the previous pick.

constexpr int PS_THREADS = 256;
constexpr int PS_KMAX = 64;

__device__ __forceinline__ bool key_less(unsigned long long da, int ja, unsigned long long db, int jb) {
    return da < db || (da == db && ja < jb);
}

__global__ __launch_bounds__(PS_THREADS) void k_pair_select(int n, const double* __restrict__ xy,
                                                             const double* __restrict__ half_long, int k,
                                                             int has_rel, int* __restrict__ out) {
    __shared__ unsigned long long s_d[PS_THREADS / 64];
    __shared__ int s_j[PS_THREADS / 64];
    __shared__ double s_nd[PS_KMAX + 1];
    __shared__ int s_nj[PS_KMAX + 1];
    const int i = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const double xi = xy[2 * i], yi = xy[2 * i + 1];
    unsigned long long prev_d = 0;
    int prev_j = -1;
    for (int r = 0; r <= k; r++) {
        unsigned long long bd = ~0ull;
        int bj = 0x7fffffff;
        for (int j = tid; j < n; j += PS_THREADS) {
            const double dx = xy[2 * j] - xi, dy = xy[2 * j + 1] - yi;
            const double dd = sqrt(dx * dx + dy * dy);
            const unsigned long long db = (unsigned long long)__double_as_longlong(dd);   // dd >= 0: bits are monotonic
            if (r > 0 && !key_less(prev_d, prev_j, db, j)) continue;
            if (key_less(db, j, bd, bj)) {
                bd = db;
                bj = j;
            }
        }
        for (int off = 32; off >= 1; off >>= 1) {
            const unsigned long long od = __shfl_xor(bd, off);
            const int oj = __shfl_xor(bj, off);
            if (key_less(od, oj, bd, bj)) {
                bd = od;
                bj = oj;
            }
        }
        if (lane == 0) {
            s_d[wv] = bd;
            s_j[wv] = bj;
        }
        __syncthreads();
        if (tid == 0) {
            for (int w = 1; w < PS_THREADS / 64; w++)
                if (key_less(s_d[w], s_j[w], bd, bj)) {
                    bd = s_d[w];
                    bj = s_j[w];
                }
            s_nd[r] = __longlong_as_double((long long)bd);
            s_nj[r] = bj;
            s_d[0] = bd;
            s_j[0] = bj;
        }
        __syncthreads();
        prev_d = s_d[0];
        prev_j = s_j[0];
        __syncthreads();
    }
    if (tid == 0) {
        // pipeline.py:463-468: pairs (min, max) of the neighbours 1..k closer than the limit
        int* o = out + (size_t)i * (k + 2) * 2;
        const double lim = has_rel ? 2.0 * half_long[i] * 0.95 : INFINITY;
        for (int r = 1; r <= k; r++) {
            const int j = s_nj[r];
            const bool keep = j < n && s_nd[r] < lim;
            o[2 * (r - 1)] = keep ? min(i, j) : -1;
            o[2 * (r - 1) + 1] = keep ? max(i, j) : -1;
        }
        // pipeline.py:470-474 temporal neighbours
        for (int off = 1; off <= 2; off++) {
            const int j = i + off;
            bool keep = false;
            if (j < n) keep = hypot(xi - xy[2 * j], yi - xy[2 * j + 1]) < 2.0 * half_long[i];
            o[2 * (k + off - 1)] = keep ? i : -1;
            o[2 * (k + off - 1) + 1] = keep ? j : -1;
        }
    }
}

}  // namespace fsx

using namespace fsx;

extern "C" {

int fsx_undistort_table_bytes(int n_img, size_t* bytes) {
    FSX_REQUIRE(n_img >= 0 && bytes, "fsx_undistort_table_bytes: bad arguments");
    *bytes = (size_t)(n_img > 0 ? n_img : 1) * sizeof(UdImage);
    return FSX_OK;
}

int fsx_undistort_batch(int n_img, const int* n_kp, const float* const* xy_dev, const double* scale,
                        const double* cam8, int iters, double* const* xy_out_dev, double* const* und_out_dev,
                        void* table_dev, size_t table_bytes, void* stream) {
    FSX_REQUIRE(n_img >= 0 && (n_img == 0 || (n_kp && xy_dev && scale && cam8 && und_out_dev && table_dev)),
                "fsx_undistort_batch: bad arguments");
    FSX_REQUIRE(iters >= 0 && iters <= 1000, "fsx_undistort_batch: iters %d out of range", iters);
    FSX_REQUIRE(table_bytes >= (size_t)n_img * sizeof(UdImage), "fsx_undistort_batch: table too small");
    if (n_img == 0) return FSX_OK;
    hipStream_t st = (hipStream_t)stream;
    UdImage* tab = new UdImage[n_img];
    int max_n = 0;
    for (int i = 0; i < n_img; i++) {
        if (n_kp[i] < 0 || (n_kp[i] > 0 && (!xy_dev[i] || !und_out_dev[i])) || !(cam8[8 * i] != 0.0)) {
            delete[] tab;
            FSX_REQUIRE(false, "fsx_undistort_batch: image %d: bad count, pointer or focal length", i);
        }
        UdImage& U = tab[i];
        U = ud_params(cam8 + 8 * i, iters);
        U.xy_in = xy_dev[i];
        U.xy_out = xy_out_dev ? xy_out_dev[i] : nullptr;
        U.und_out = und_out_dev[i];
        U.n = n_kp[i];
        U.sx = scale[2 * i];
        U.sy = scale[2 * i + 1];
        max_n = n_kp[i] > max_n ? n_kp[i] : max_n;
    }
    int rc = upload_async(table_dev, tab, (size_t)n_img * sizeof(UdImage), st);
    delete[] tab;
    if (rc) return rc;
    if (max_n == 0) return FSX_OK;
    FSX_REQUIRE(n_img <= 65535, "fsx_undistort_batch: at most 65535 images per call");
    hipLaunchKernelGGL(k_undistort, dim3((max_n + 255) / 256, n_img), dim3(256), 0, st, (const UdImage*)table_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_undistort_f64(int n, const double* uv_dev, const double* cam8, int iters, double* out_dev, void* stream) {
    FSX_REQUIRE(n >= 0 && cam8 && (n == 0 || (uv_dev && out_dev)), "fsx_undistort_f64: bad arguments");
    FSX_REQUIRE(iters >= 0 && iters <= 1000, "fsx_undistort_f64: iters %d out of range", iters);
    FSX_REQUIRE(cam8[0] != 0.0, "fsx_undistort_f64: focal length is zero");
    if (n == 0) return FSX_OK;
    UdImage U = ud_params(cam8, iters);
    hipLaunchKernelGGL(k_undistort_f64, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, uv_dev, U,
                       out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_ground_predict(int n, const double* u_dev, const double* cams, double* pred_dev, uint8_t* ok_dev,
                       void* stream) {
    FSX_REQUIRE(n >= 0 && cams && (n == 0 || (u_dev && pred_dev && ok_dev)), "fsx_ground_predict: bad arguments");
    if (n == 0) return FSX_OK;
    GroundCams P;
    memcpy(P.Kinv_i, cams, 9 * sizeof(double));
    memcpy(P.R_i, cams + 9, 9 * sizeof(double));
    memcpy(P.C_i, cams + 18, 3 * sizeof(double));
    P.zg_i = cams[21];
    memcpy(P.R_j, cams + 22, 9 * sizeof(double));
    memcpy(P.K_j, cams + 31, 9 * sizeof(double));
    memcpy(P.C_j, cams + 40, 3 * sizeof(double));
    hipLaunchKernelGGL(k_ground_predict, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, n, u_dev, P,
                       pred_dev, ok_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_pair_select(int n, const double* xy_dev, const double* half_long_dev, int k, int has_rel, int* out_dev,
                    void* stream) {
    FSX_REQUIRE(n >= 0 && k >= 0 && k <= PS_KMAX && k <= (n > 0 ? n - 1 : 0), "fsx_pair_select: bad n=%d k=%d", n, k);
    FSX_REQUIRE(n == 0 || (xy_dev && half_long_dev && out_dev), "fsx_pair_select: null pointer");
    if (n == 0) return FSX_OK;
    hipLaunchKernelGGL(k_pair_select, dim3(n), dim3(PS_THREADS), 0, (hipStream_t)stream, n, xy_dev, half_long_dev, k,
                       has_rel, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}


int fsx_prior_keep_batch(int n_pair, const long long* off_dev, const double* ua_dev, const double* ub_dev,
                         const double* cams_dev, const double* tol_dev, int max_rows, uint8_t* keep_dev,
                         void* stream) {
    FSX_REQUIRE(n_pair >= 0 && n_pair <= 65535 && max_rows >= 0, "fsx_prior_keep_batch: bad sizes");
    if (n_pair == 0 || max_rows == 0) return FSX_OK;
    FSX_REQUIRE(off_dev && ua_dev && ub_dev && cams_dev && tol_dev && keep_dev, "fsx_prior_keep_batch: null pointer");
    hipLaunchKernelGGL(k_prior_keep, dim3((unsigned)((max_rows + 255) / 256), (unsigned)n_pair), dim3(256), 0,
                       (hipStream_t)stream, off_dev, ua_dev, ub_dev, (const GroundCams*)cams_dev, tol_dev, keep_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
