// SIFT detector/descriptor kernels for gfx950 (CDNA4): Gaussian scale space, DoG extrema,
// sub-pixel refinement, orientation histograms, deterministic keypoint ordering and 128-d
// descriptors.  Arithmetic reproduces the reference's CPU path (app/planesite/gsift.py with
// numpy/scipy dtypes): float32 where numpy is float32 (compiled with -ffp-contract=off, so no
// fused multiply-add), float64 where numpy promotes (scipy's correlate1d accumulation,
// bincount, the descriptor's sample geometry).  Histograms accumulate in 64-bit fixed point
// with LDS integer atomics, so every run gives the same bits.
#include "fsx_common.h"
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

namespace fsx {

// ------------------------------------------------------------------------------------------
// Gaussian taps (gsift.py:52-72).  Host side, float64 like numpy, normalised with numpy's
// pairwise summation order, cast to float32.

static double np_sum_d(const double* a, int n) {
    if (n < 8) {
        double r = 0.0;
        for (int i = 0; i < n; i++) r += a[i];
        return r;
    }
    double r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int i;
    for (i = 8; i < n - (n % 8); i += 8)
        for (int j = 0; j < 8; j++) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
}

static int gauss_taps(double sigma, float* out) {
    int k = ((int)rint(sigma * 8.0 + 1.0)) | 1;
    if (k > 63) return -1;
    double g[64];
    for (int i = 0; i < k; i++) {
        double x = (double)i - (k - 1) / 2.0;
        g[i] = exp(-x * x / (2.0 * sigma * sigma));
    }
    double s = np_sum_d(g, k);
    for (int i = 0; i < k; i++) out[i] = (float)(g[i] / s);
    return k;
}

static void layer_sigmas(double* sig6) {
    const double SIG = 1.6, INIT = 0.5;
    double k = pow(2.0, 1.0 / N_LAYERS);
    sig6[0] = sqrt(fmax(SIG * SIG - (2 * INIT) * (2 * INIT), 0.01));
    for (int i = 1; i < N_LAYERS + 3; i++) {
        double prev = pow(k, i - 1) * SIG;
        double total = prev * k;
        sig6[i] = sqrt(total * total - prev * prev);
    }
}

// Half kernel by distance from the centre: k[0] = centre, k[j] = left tap at distance j.
struct HalfTaps {
    double k[16];
};

struct TapSet {
    int r2[6];
    HalfTaps h[6];
    float full[6][64];
    int len[6];
};

static const TapSet& taps() {
    static TapSet T;
    static bool init = false;
    if (!init) {
        double sig[6];
        layer_sigmas(sig);
        for (int i = 0; i < 6; i++) {
            T.len[i] = gauss_taps(sig[i], T.full[i]);
            T.r2[i] = T.len[i] / 2;
            for (int j = 0; j <= T.r2[i]; j++) T.h[i].k[j] = (double)T.full[i][T.r2[i] - j];
        }
        init = true;
    }
    return T;
}

// gsift.py:158-172: octave count from the input's smaller side, first octave = 2x upsample,
// each next octave = ceil halves, stop below 2*BORDER+4.
int sift_plan(int H, int W, SiftPlan* p) {
    memset(p, 0, sizeof(*p));
    p->H = H;
    p->W = W;
    int mn = H < W ? H : W;
    int n_oct = (int)rint(log2((double)mn) - 2.0) + 1;
    if (n_oct < 1) n_oct = 1;
    int r = 2 * H, c = 2 * W, n = 0;
    for (int o = 0; o < n_oct && o < MAXO; o++) {
        if ((r < c ? r : c) < 2 * BORDER + 4) break;
        p->R[n] = r;
        p->C[n] = c;
        p->sumP += (long long)r * c;
        n++;
        r = (r + 1) / 2;
        c = (c + 1) / 2;
    }
    p->n_oct = n;
    p->P0 = (long long)p->R[0] * p->C[0];
    return n;
}

// ------------------------------------------------------------------------------------------
// Pyramid: one LDS-staged 2-D tile kernel per (octave, layer), batched over images.
// Vertical pass (axis 0) then horizontal pass (axis 1), each rounded to float32 like
// scipy.ndimage.correlate1d; float64 accumulation: centre tap first, then symmetric pairs
// from the outermost inwards (scipy ni_filters.c, symmetric branch).  Mirror border =
// reflect-101.  The epilogue writes the new gauss layer and, for layer 3, the next octave's
// base (every second pixel).  The DoG atlas is never stored: its readers form
// G[l+1] - G[l] in float32 (the reference's DF values bit for bit), which saves a 4-byte
// write and a re-read of the previous layer per pixel and layer.


// global-address-space view of a generic pointer: lets the compiler emit global_load (vector
// memory counter only) instead of flat_load for data read through pointers held in structs
template <typename T>
__device__ __forceinline__ const __attribute__((address_space(1))) T* gptr(const T* p) {
    return (const __attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ int reflect101(int p, int n) {
    while (p < 0 || p >= n) p = p < 0 ? -p : 2 * n - 2 - p;
    return p;
}

// value of the 2x bilinear upsample of u8 * float32(1/255) at (y, x) (gsift.py:90-102, 160)
__device__ __forceinline__ float up2_value(const uint8_t* __restrict__ img, int H, int W, int ld, int y, int x) {
    const float s255 = (float)(1.0 / 255.0);
    int iy = y >> 1, ix = x >> 1;
    int ny = (y & 1) ? min(iy + 1, H - 1) : max(iy - 1, 0);
    int nx = (x & 1) ? min(ix + 1, W - 1) : max(ix - 1, 0);
    const __attribute__((address_space(1))) uint8_t* g = gptr(img);
    float a_ii = (float)g[(size_t)iy * ld + ix] * s255;
    float a_ni = (float)g[(size_t)ny * ld + ix] * s255;
    float a_in = (float)g[(size_t)iy * ld + nx] * s255;
    float a_nn = (float)g[(size_t)ny * ld + nx] * s255;
    float t_i = 0.75f * a_ii + 0.25f * a_ni;
    float t_n = 0.75f * a_in + 0.25f * a_nn;
    return 0.75f * t_i + 0.25f * t_n;
}

// Tile: 256 input columns, TWo output columns (the largest multiple of 16 <= 256 - 2*R2: 240
// for R2 <= 8, 224 above; tiles start on 64-byte boundaries of their row, so where rows are
// 64-byte aligned, as in octave 0 of most photo widths, each row segment is written as whole
// 64-byte sectors, and the vertical pass computes at most 16 + 2*R2 columns it does not store) x TH output rows
// (the row strip: 8, 16, 24 or 32, the tuning key TUNE_BLUR_ROWS), 256 threads.
// Vertical pass: a thread owns one input column and the TH output rows; it streams the
// TH + 2*R2 input values of its column once from global memory (coalesced across the wave),
// converts each to float64 once and accumulates from registers.  The float32-rounded result
// goes to an LDS mid-tile (row pitch 257 floats).  Horizontal pass: a thread owns CB (about
// TWo * TH / 256) consecutive outputs of one row, reading CB + 2*R2 mid values once.  Epilogue: lanes along
// x, coalesced stores of the gauss layer (the next launch's input).  A taller strip reads
// fewer halo rows per output (TH + 2*R2 per TH) but needs more registers and LDS per block.
constexpr int BL_SW = 256;
constexpr int BL_NT = 256;

template <int R2>
constexpr int blur_tw() { return (BL_SW - 2 * R2) & ~15; }

// Workgroups are dispatched to the 8 XCDs round-robin by linear id; remap so that XCD k gets a
// contiguous range of row-major tiles (a band of tile rows), keeping vertically adjacent tiles,
// whose halos overlap, on the same L2.  A bijection for any grid size.
__device__ __forceinline__ void xcd_tile(int& tx, int& ty, int& tz) {
    const unsigned nx = gridDim.x, ny = gridDim.y, nz = gridDim.z;
    const unsigned N = nx * ny * nz;
    const unsigned L = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const unsigned q = N / 8, r = N % 8, k = L % 8, j = L / 8;
    const unsigned T = k * q + (k < r ? k : r) + j;
    tx = (int)(T % nx);
    ty = (int)((T / nx) % ny);
    tz = (int)(T / (nx * ny));
}

// One output of the symmetric correlation, rounded to float32 exactly as the reference's
// float64 loop rounds it: centre product first, then acc += (w[-j] + w[+j]) * k[j] for
// j = R2 .. 1, each product and sum rounded (gsift.py:119-147 via scipy's correlate1d).
// Fast form: the products are fused (acc = fma(s, k, acc)).  Both forms evaluate the same
// non-negative terms (image data and Gaussian taps are >= 0, the pair sums are identical), so
// each lies within (2*R2 + 1) * 2^-53 * S of the exact sum S; they differ by < 41 * 2^-53 * S
// < 2^-47 * acc, i.e. by less than 64 float64 ulps of acc.  Round-to-nearest float32 only
// jumps at the midpoints between float32 values, which are the float64 values whose 29
// mantissa bits below float32 precision read 0x10000000; if acc's low 29 bits are more than
// 256 ulps away from that pattern, no midpoint separates acc from the reference's value and
// both round to the same float32 (rounding is monotone).  Otherwise (~1 output in 2^20), and
// where the float32 result is below 2^-120 (near float32's subnormal range, 2^-126, whose
// rounding points differ) but not 0, the reference's unfused order is evaluated.
// The test costs four instructions: low29 - (0x10000000 - 256) <= 512 is evaluated as
// (lo << 3) - (0x0FFFFF00 << 3) <= 4096 in 32 bits (the shift drops the bits above the low 29,
// one shift-add), and 0 < r < 2^-120 as bits(r) - 1 < 0x037FFFFF on the float32 result r
// (an acc just below 2^-120 that rounds up to it is >= 2^-126, where the midpoint argument holds).
template <int R2>
__device__ __forceinline__ float blur_point(const double* w, const HalfTaps& tp) {
    double acc = w[R2] * tp.k[0];
#pragma unroll
    for (int j = R2; j >= 1; --j) acc = __builtin_fma(w[R2 - j] + w[R2 + j], tp.k[j], acc);
    const unsigned lo = (unsigned)__double_as_longlong(acc);
    const float r = (float)acc;
    const bool near_mid = (lo << 3) - (0x0FFFFF00u << 3) <= 4096u;
    const bool tiny = __float_as_uint(r) - 1u < 0x037FFFFFu;       // acc >= 0: r >= 0
    if (__builtin_expect(!(near_mid || tiny), 1)) return r;
    double ex = w[R2] * tp.k[0];
#pragma unroll
    for (int j = R2; j >= 1; --j) ex += (w[R2 - j] + w[R2 + j]) * tp.k[j];
    return (float)ex;
}

// blur_point for a thread's group of outputs: each output's fused result is rounded and its two
// check quantities folded into running minima (mid: (lo << 3) - (0x0FFFFF00 << 3), tiny:
// bits(r) - 1, both unsigned), so the group costs one branch instead of one per output; when
// any output of the group fails the check (mid <= 4096 or tiny < 0x037FFFFF) the group is
// evaluated again output by output with blur_point.  Either way every output equals
// blur_point's.
template <int R2>
__device__ __forceinline__ double blur_acc(const double* w, const HalfTaps& tp) {
    double acc = w[R2] * tp.k[0];
#pragma unroll
    for (int j = R2; j >= 1; --j) acc = __builtin_fma(w[R2 - j] + w[R2 + j], tp.k[j], acc);
    return acc;
}

__device__ __forceinline__ float blur_round(double acc, unsigned& mid, unsigned& tiny) {
    const unsigned lo = (unsigned)__double_as_longlong(acc);
    const float r = (float)acc;
    mid = min(mid, (lo << 3) - (0x0FFFFF00u << 3));
    tiny = min(tiny, __float_as_uint(r) - 1u);
    return r;
}

// (the empty asm with a memory clobber keeps the group's stores ahead of the branch: without it
// the compiler sinks them below, since the per-output path overwrites the same words, and holds
// every rounded value of the group in registers until then)
__device__ __forceinline__ bool blur_group_unsure(unsigned mid, unsigned tiny) {
    asm volatile("" ::: "memory");
    return mid <= 4096u || tiny < 0x037FFFFFu;
}

// the per-output path's copy of a window, passed through an opaque move so that none of its
// pair sums is shared with the group's fast evaluation (shared sums would stay live across the
// branch, all of them at once)
template <int N>
__device__ __forceinline__ void blur_launder(const double* w, double* out) {
#pragma unroll
    for (int k = 0; k < N; k++) {
        double x = w[k];
        asm volatile("" : "+v"(x));
        out[k] = x;
    }
}

template <int R2, bool BASE, int BL_TH>
__global__ __launch_bounds__(BL_NT) void k_blur(const SiftJob* __restrict__ jobs, int o, int layer, HalfTaps tp) {
    constexpr int TWo = blur_tw<R2>();
    constexpr int NW = BL_TH + 2 * R2;        // vertical window
    // horizontal tasks: CB consecutive outputs of one row each, at most 256 / TH tasks per row,
    // so that the TH x TWo outputs make at most one task per thread (14 outputs per task for
    // TWo = 224 at TH = 16: every lane of the four waves works in both passes)
    constexpr int NGMAX = BL_NT / BL_TH;
    constexpr int CB = (TWo + NGMAX - 1) / NGMAX;
    constexpr int HW = CB + 2 * R2;           // horizontal window
    constexpr int NG = (TWo + CB - 1) / CB;
    constexpr int NWV = BL_NT / 64;           // waves
    constexpr int EPR = BL_TH / NWV;          // epilogue rows per wave
    constexpr int EPC = (TWo + 63) / 64;      // epilogue column steps per lane
    static_assert(BL_SW == BL_NT && BL_TH % NWV == 0 && BL_TH % 2 == 0, "one input column per thread, even rows");
    __shared__ float s_mid[BL_TH][BL_SW + 1];
    __shared__ float s_out[BL_TH][TWo + 1];   // odd pitch: the horizontal pass writes conflict-free;
                                              // 31.9 KB in all at 16 rows: 5 blocks per CU where VGPRs allow
    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const SiftJob& J = jobs[tz];
    if (o >= J.n_oct) return;
    const int R = J.oct[o].R, C = J.oct[o].C;
    const int x0 = tx * TWo, y0 = ty * BL_TH;
    if (x0 >= C || y0 >= R) return;
    const long long P = (long long)R * C;
    const float* in = nullptr;
    if (!BASE) {
        if (layer == 1) in = J.G0 + J.oct[o].g0base;
        else if (layer <= 4) in = J.G123 + J.oct[o].gbase + (long long)(layer - 2) * P;
        else in = J.G45 + J.oct[o].g45base;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    {
        const int c = tid;
        const int gx = reflect101(x0 - R2 + c, C);
        const int ybase = y0 - R2;
        // the thread's input column over the window rows, as float64 (called again by the
        // group's per-output path, so that the window is not held across the group)
        auto load_window = [&](double (&w)[NW]) {
            float f[NW];
            // base layer: ybase = TH*ty - R2 has the parity of R2, so in the interior the source
            // rows of each upsampled row are fixed offsets from rb = (ybase >> 1) - 1; the thread
            // loads its two source columns over those NI rows once (the same float32 operations
            // as up2_value, no clamping needed inside [rb, rb + NI))
            constexpr int PAR = R2 & 1;
            constexpr int NI = ((PAR + NW - 1) >> 1) + 3;
            const int rb = (ybase >> 1) - 1;
            if (BASE && ybase >= 0 && rb >= 0 && rb + NI <= J.H) {
                const float s255 = (float)(1.0 / 255.0);
                const int ix = gx >> 1;
                const int nx = (gx & 1) ? min(ix + 1, J.W - 1) : max(ix - 1, 0);
                const __attribute__((address_space(1))) uint8_t* src = gptr(J.img) + (size_t)rb * J.ld;
                float bi[NI], bn[NI];
#pragma unroll
                for (int r = 0; r < NI; r++) {
                    bi[r] = (float)src[(size_t)r * J.ld + ix] * s255;
                    bn[r] = (float)src[(size_t)r * J.ld + nx] * s255;
                }
#pragma unroll
                for (int k = 0; k < NW; k++) {
                    const int iy = ((PAR + k) >> 1) + 1;
                    const int ny = ((PAR + k) & 1) ? iy + 1 : iy - 1;
                    const float t_i = 0.75f * bi[iy] + 0.25f * bi[ny];
                    const float t_n = 0.75f * bn[iy] + 0.25f * bn[ny];
                    f[k] = 0.75f * t_i + 0.25f * t_n;
                }
            } else if (ybase >= 0 && ybase + NW <= R) {
                // interior rows (block-uniform test): direct addressing, no reflection
                if (BASE) {
#pragma unroll
                    for (int k = 0; k < NW; k++) f[k] = up2_value(J.img, J.H, J.W, J.ld, ybase + k, gx);
                } else {
                    const __attribute__((address_space(1))) float* col = gptr(in) + (size_t)ybase * C + gx;
#pragma unroll
                    for (int k = 0; k < NW; k++) f[k] = col[(size_t)k * C];
                }
            } else {
#pragma unroll
                for (int k = 0; k < NW; k++) {
                    const int gy = reflect101(ybase + k, R);
                    f[k] = BASE ? up2_value(J.img, J.H, J.W, J.ld, gy, gx) : gptr(in)[(size_t)gy * C + gx];
                }
            }
#pragma unroll
            for (int k = 0; k < NW; k++) w[k] = (double)f[k];
        };
        double w[NW];
        load_window(w);
        unsigned mid = ~0u, tiny = ~0u;
#pragma unroll
        for (int i = 0; i < BL_TH; i++) s_mid[i][c] = blur_round(blur_acc<R2>(w + i, tp), mid, tiny);
        if (__builtin_expect(blur_group_unsure(mid, tiny), 0)) {
            double wl[NW];
            load_window(wl);
#pragma unroll
            for (int i = 0; i < BL_TH; i++) s_mid[i][c] = blur_point<R2>(wl + i, tp);
        }
    }
    __syncthreads();
    for (int task = tid; task < BL_TH * NG; task += BL_NT) {
        const int ly = task % BL_TH, cg = task / BL_TH;
        const int cbase = cg * CB;
        float fv[HW];
#pragma unroll
        for (int k = 0; k < HW; k++) fv[k] = (cbase + k < BL_SW) ? s_mid[ly][cbase + k] : 0.0f;
        double v[HW];
#pragma unroll
        for (int k = 0; k < HW; k++) v[k] = (double)fv[k];
        unsigned mid = ~0u, tiny = ~0u;
#pragma unroll
        for (int j = 0; j < CB; j++) {
            if (cbase + j >= TWo) continue;
            s_out[ly][cbase + j] = blur_round(blur_acc<R2>(v + j, tp), mid, tiny);
        }
        if (__builtin_expect(blur_group_unsure(mid, tiny), 0)) {
            // the window again from LDS (not held across the group)
            double vl[HW];
#pragma unroll
            for (int k = 0; k < HW; k++) vl[k] = (double)((cbase + k < BL_SW) ? s_mid[ly][cbase + k] : 0.0f);
            blur_launder<HW>(vl, vl);
#pragma unroll
            for (int j = 0; j < CB; j++) {
                if (cbase + j >= TWo) continue;
                s_out[ly][cbase + j] = blur_point<R2>(vl + j, tp);
            }
        }
    }
    __syncthreads();
    // epilogue with lanes along x: coalesced gauss / downsample stores (the DoG is never
    // stored: its readers form G[l+1] - G[l], the same float32 subtraction)
    float* outG = nullptr;
    float* down = nullptr;
    int R2n = 0, C2n = 0;
    if (BASE) {
        outG = J.G0 + J.oct[0].g0base;
    } else {
        if (layer <= 3) outG = J.G123 + J.oct[o].gbase + (long long)(layer - 1) * P;
        else outG = J.G45 + J.oct[o].g45base + (long long)(layer - 4) * P;
        if (layer == 3 && o + 1 < J.n_oct) {
            down = J.G0 + J.oct[o + 1].g0base;
            R2n = J.oct[o + 1].R;
            C2n = J.oct[o + 1].C;
        }
    }
#pragma unroll
    for (int r = 0; r < EPR; r++) {
        const int ly = wv + r * NWV, gy = y0 + ly;
#pragma unroll
        for (int cc = 0; cc < EPC; cc++) {
            const int lx = lane + 64 * cc, gx = x0 + lx;
            if (gy < R && lx < TWo && gx < C) {
                const float g = s_out[ly][lx];
                const size_t p = (size_t)gy * C + gx;
                outG[p] = g;
                if (down && !(gy & 1) && !(gx & 1)) {
                    const int ry = gy >> 1, rx = gx >> 1;
                    if (ry < R2n && rx < C2n) down[(size_t)ry * C2n + rx] = g;
                }
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Extrema (gsift.py:293-319, 633-659): non-strict 3x3x3 max/min of DoG layers 1..3 with the
// prefilter |v| > 0.8*0.5*contrast/3, outside a 5-pixel border.  Candidates are appended in
// any order as gauss-atlas indices gbase[o] + (l-1)*R*C + r*C + c (ascending index = the
// reference's candidate order; the final sort key carries it).

constexpr int EX_W = 62;     // output columns per wave (lanes 1..62; lanes 0 and 63 are halo)
constexpr int EX_H = 16;     // output rows per block (a sweep of 16..128: 16 is fastest)

__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ int wave_append(bool f, int* counter) {
    unsigned long long m = __ballot(f);
    int cnt = __popcll(m);
    int base = 0;
    const int lane = lane_id();
    const int leader = __ffsll((long long)m) - 1;
    if (cnt && lane == leader) base = atomicAdd(counter, cnt);
    base = __shfl(base, leader < 0 ? 0 : leader);
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

// Finite differences of gsift.py:274-290 at one DoG voxel.  derivs_from takes the 19 samples
// of the 3x3x3 cube (x / y / s = column / row / layer, p / m = +1 / -1); the extrema kernel
// feeds it from registers and derivs_at from the gauss planes, so both give the same bits.
struct Derivs {
    float v, dx, dy, ds, dxx, dyy, dss, dxy, dxs, dys;
};
constexpr int DER_N = 10;   // floats of a Derivs record in the candidate side array

__device__ __forceinline__ Derivs derivs_from(float v, float xp, float xm, float yp, float ym, float sp, float sm,
                                              float ypxp, float ypxm, float ymxp, float ymxm, float spxp,
                                              float spxm, float smxp, float smxm, float spyp, float spym,
                                              float smyp, float smym) {
    Derivs d;
    d.v = v;
    d.dx = (xp - xm) * 0.5f;
    d.dy = (yp - ym) * 0.5f;
    d.ds = (sp - sm) * 0.5f;
    d.dxx = (xp + xm) - 2.0f * v;
    d.dyy = (yp + ym) - 2.0f * v;
    d.dss = (sp + sm) - 2.0f * v;
    d.dxy = (((ypxp - ypxm) - ymxp) + ymxm) * 0.25f;
    d.dxs = (((spxp - spxm) - smxp) + smxm) * 0.25f;
    d.dys = (((spyp - spym) - smyp) + smym) * 0.25f;
    return d;
}

// gauss plane l (0..5) of octave o
__device__ __forceinline__ const float* gauss_plane(const SiftJob& J, int o, int l, long long P) {
    return l == 0 ? J.G0 + J.oct[o].g0base
                  : (l <= 3 ? J.G123 + J.oct[o].gbase + (long long)(l - 1) * P
                            : J.G45 + J.oct[o].g45base + (long long)(l - 4) * P);
}

// Derivatives at DoG layer l (1..3), plane offset k = r*C + c, from gauss planes l-1 .. l+2
// (DoG layer m = G[m+1] - G[m]).
__device__ __forceinline__ Derivs derivs_at(const SiftJob& J, int o, int l, long long k, long long C, long long P) {
    const float* __restrict__ g0 = gauss_plane(J, o, l - 1, P);
    const float* __restrict__ g1 = gauss_plane(J, o, l, P);
    const float* __restrict__ g2 = gauss_plane(J, o, l + 1, P);
    const float* __restrict__ g3 = gauss_plane(J, o, l + 2, P);
    auto dm = [&](long long q) { return g1[q] - g0[q]; };   // DoG layer l-1
    auto d0 = [&](long long q) { return g2[q] - g1[q]; };   // DoG layer l
    auto dp = [&](long long q) { return g3[q] - g2[q]; };   // DoG layer l+1
    return derivs_from(d0(k), d0(k + 1), d0(k - 1), d0(k + C), d0(k - C), dp(k), dm(k), d0(k + C + 1), d0(k + C - 1),
                       d0(k - C + 1), d0(k - C - 1), dp(k + 1), dp(k - 1), dm(k + 1), dm(k - 1), dp(k + C), dp(k - C),
                       dm(k + C), dm(k - C));
}

// Streaming-column extrema: each wave owns 62 columns of one octave and walks down EX_H
// rows; per row it loads the 6 gauss layers (coalesced), forms the 5 DoG layers, takes the
// horizontal max/min over
// (x-1, x, x+1) with lane shuffles and keeps a 3-row ring, so the 3x3x3 max/min of a voxel
// costs a few max/min operations.  v >= max27 is the reference's non-strict test (no NaN).
// Candidates are staged per wave in LDS and appended EX_BUF at a time with one atomic on the
// image's counter (an atomic per row with candidates, all on one address, cost a third of the
// kernel).  The kernel also keeps the raw 3-row ring, so it hands each candidate's first
// refinement derivatives to k_refine (three of four candidates converge without moving, and
// then refinement gathers nothing).
constexpr int EX_BUF = 64;
static_assert(EX_BUF >= EX_W, "one layer's candidates of a row always fit an empty buffer");

__global__ __launch_bounds__(256) void k_extrema(const SiftJob* __restrict__ jobs, int o) {
    __shared__ uint32_t s_buf[4][EX_BUF];
    __shared__ float s_der[4][EX_BUF * DER_N];
    int tx, ty, tz;
    xcd_tile(tx, ty, tz);
    const SiftJob& J = jobs[tz];
    if (o >= J.n_oct) return;
    const int R = J.oct[o].R, C = J.oct[o].C;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    uint32_t* buf = s_buf[wv];
    float* der = s_der[wv];
    int nbuf = 0;   // staged candidates of this wave (wave-uniform)
    // waves of a block may leave early, so the buffer is ordered with wave-level barriers
    auto flush = [&]() {
        if (nbuf == 0) return;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        int base = 0;
        if (lane == 0) base = atomicAdd(&J.counters[0], nbuf);
        base = __shfl(base, 0);
        for (int i = lane; i < nbuf; i += 64) {
            if (base + i < J.cand_cap) J.cand[base + i] = buf[i];
            else atomicOr(&J.counters[3], 1);
        }
        const int nder = min(nbuf, J.cand_cap - base) * DER_N;   // <= 0 once the array is full
        for (int i = lane; i < nder; i += 64) J.cand_der[(size_t)base * DER_N + i] = der[i];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        nbuf = 0;
    };
    const int xs = (tx * 4 + wv) * EX_W;                   // first output column of this wave
    const int y0 = ty * EX_H;
    if (xs >= C || y0 >= R) return;                        // wave-uniform (no block barrier below)
    const int x = xs - 1 + lane;
    const int xc = min(max(x, 0), C - 1);
    const long long P = (long long)R * C;
    // the six gauss planes of this octave; DoG layer l = G[l+1] - G[l] (float32, as the
    // reference's DF atlas holds it)
    const __attribute__((address_space(1))) float* Gp[6] = {
        gptr(J.G0 + J.oct[o].g0base + xc), gptr(J.G123 + J.oct[o].gbase + xc),
        gptr(J.G123 + J.oct[o].gbase + P + xc), gptr(J.G123 + J.oct[o].gbase + 2 * P + xc),
        gptr(J.G45 + J.oct[o].g45base + xc), gptr(J.G45 + J.oct[o].g45base + P + xc)};
    const float pthr = J.pthr;
    const bool col_ok = lane >= 1 && lane <= EX_W && x >= BORDER && x < C - BORDER;
    float hmx[3][5], hmn[3][5];
    float prev[2][5];   // the last two rows of the previous 4-row block
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int l = 0; l < 5; l++) {
            hmx[a][l] = 0.f;
            hmn[a][l] = 0.f;
        }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int l = 0; l < 5; l++) prev[a][l] = 0.f;
    const int yend = min(y0 + EX_H, R);
    // rows are fetched 4 at a time, one block of 4 ahead of the one being processed
    float nxt[4][6];
#pragma unroll
    for (int r = 0; r < 4; r++) {
        const int yc = min(max(y0 - 1 + r, 0), R - 1);
#pragma unroll
        for (int l = 0; l < 6; l++) nxt[r][l] = Gp[l][(long long)yc * C];
    }
    for (int yb = y0 - 1; yb <= yend; yb += 4) {
        float cg[4][6];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int l = 0; l < 6; l++) cg[r][l] = nxt[r][l];
        if (yb + 4 <= yend) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int yc = min(max(yb + 4 + r, 0), R - 1);
#pragma unroll
                for (int l = 0; l < 6; l++) nxt[r][l] = Gp[l][(long long)yc * C];
            }
        }
        // DoG of this block's rows, formed after the next block's loads are in flight
        float cur[4][5];
#pragma unroll
        for (int r = 0; r < 4; r++)
#pragma unroll
            for (int l = 0; l < 5; l++) cur[r][l] = cg[r][l + 1] - cg[r][l];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int yy = yb + r;
            if (yy > yend) break;                           // wave-uniform
#pragma unroll
            for (int l = 0; l < 5; l++) {
                const float v = cur[r][l];
                // neighbours x-1 / x+1 by whole-wave DPP shifts (max/min of the three is
                // symmetric, the halo lanes 0 and 63 never produce an output)
                const float lf = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
                const float rt = __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
                hmx[0][l] = hmx[1][l];
                hmx[1][l] = hmx[2][l];
                hmx[2][l] = fmaxf(fmaxf(lf, v), rt);
                hmn[0][l] = hmn[1][l];
                hmn[1][l] = hmn[2][l];
                hmn[2][l] = fminf(fminf(lf, v), rt);
            }
            // raw[a][l]: layer l of rows y-1, y, y+1 (= yy-2 .. yy) of this column
            float raw[3][5];
#pragma unroll
            for (int a = 0; a < 3; a++)
#pragma unroll
                for (int l = 0; l < 5; l++) raw[a][l] = r - 2 + a >= 0 ? cur[(r - 2 + a) & 3][l] : prev[(r + a) & 1][l];
            const int y = yy - 1;                           // the ring's middle row
            if (yy < y0 + 1) continue;                      // wave-uniform
            const bool row_ok = y >= BORDER && y < R - BORDER;
            float vm[5], vn[5];
#pragma unroll
            for (int l = 0; l < 5; l++) {
                vm[l] = fmaxf(fmaxf(hmx[0][l], hmx[1][l]), hmx[2][l]);
                vn[l] = fminf(fminf(hmn[0][l], hmn[1][l]), hmn[2][l]);
            }
            bool f[3];
#pragma unroll
            for (int l = 1; l <= N_LAYERS; l++) {
                const float c = raw[1][l];
                const float mx = fmaxf(fmaxf(vm[l - 1], vm[l]), vm[l + 1]);
                const float mn = fminf(fminf(vn[l - 1], vn[l]), vn[l + 1]);
                f[l - 1] = col_ok && row_ok && ((c >= mx && c > pthr) || (c <= mn && c < -pthr));
            }
            const unsigned long long b0 = __ballot(f[0]), b1 = __ballot(f[1]), b2 = __ballot(f[2]);
            const int n0 = __popcll(b0), n1 = __popcll(b1), n2 = __popcll(b2);
            if (n0 + n1 + n2 == 0) continue;                // wave-uniform
            const unsigned long long below = (1ull << lane) - 1ull;
#pragma unroll
            for (int l = 1; l <= N_LAYERS; l++) {
                const unsigned long long bl = l == 1 ? b0 : (l == 2 ? b1 : b2);
                if (bl == 0) continue;                      // wave-uniform
                const int nl = __popcll(bl);
                if (nbuf + nl > EX_BUF) flush();
                // the 3x3x3 cube around (l, y, x): rows from the ring, columns x -/+ 1 by DPP
                // (all lanes take part; lanes 0 and 63 never hold a candidate)
                auto xm = [](float v) {
                    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
                };
                auto xp = [](float v) {
                    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
                };
                const Derivs d = derivs_from(raw[1][l], xp(raw[1][l]), xm(raw[1][l]), raw[2][l], raw[0][l],
                                             raw[1][l + 1], raw[1][l - 1], xp(raw[2][l]), xm(raw[2][l]),
                                             xp(raw[0][l]), xm(raw[0][l]), xp(raw[1][l + 1]), xm(raw[1][l + 1]),
                                             xp(raw[1][l - 1]), xm(raw[1][l - 1]), raw[2][l + 1], raw[0][l + 1],
                                             raw[2][l - 1], raw[0][l - 1]);
                if (f[l - 1]) {
                    const int p = nbuf + (int)__popcll(bl & below);
                    buf[p] = (uint32_t)(J.oct[o].gbase + (long long)(l - 1) * P + (long long)y * C + x);
                    float* q = der + p * DER_N;
                    q[0] = d.v; q[1] = d.dx; q[2] = d.dy; q[3] = d.ds; q[4] = d.dxx;
                    q[5] = d.dyy; q[6] = d.dss; q[7] = d.dxy; q[8] = d.dxs; q[9] = d.dys;
                }
                nbuf += nl;
            }
        }
#pragma unroll
        for (int l = 0; l < 5; l++) {
            prev[0][l] = cur[2][l];
            prev[1][l] = cur[3][l];
        }
    }
    flush();
}

// ------------------------------------------------------------------------------------------
// Refinement (gsift.py:274-290 derivatives, 182-198 adjugate solve, 322-369 iteration and
// the contrast / edge tests).  One thread per candidate.

__global__ __launch_bounds__(256) void k_refine(const SiftJob* __restrict__ jobs) {
    const SiftJob& J = jobs[blockIdx.z];
    const int n = min(J.counters[0], J.cand_cap);
    // block-uniform loop so that every lane of a wave reaches the ballot in wave_append
    for (int base = blockIdx.x * 256; base < n; base += gridDim.x * 256) {
        const int i = base + threadIdx.x;
        bool keep = false;
        int o = 0, l = 0, r = 0, c = 0;
        float X0 = 0.f, X1 = 0.f, X2 = 0.f, resp = 0.f;
        uint32_t g = 0;
        if (i < n) {
            g = J.cand[i];
            while (o + 1 < J.n_oct && (long long)g >= J.oct[o + 1].gbase) o++;
            const int R = J.oct[o].R, C = J.oct[o].C;
            const long long RC = (long long)R * C;
            long long loc = (long long)g - J.oct[o].gbase;
            l = (int)(loc / RC) + 1;
            long long rem = loc - (long long)(l - 1) * RC;
            r = (int)(rem / C);
            c = (int)(rem - (long long)r * C);
            bool good = false;
            // derivatives at the candidate itself come from the extrema kernel
            const float2* q = reinterpret_cast<const float2*>(J.cand_der) + (size_t)i * (DER_N / 2);
            const float2 q0 = q[0], q1 = q[1], q2 = q[2], q3 = q[3], q4 = q[4];
            Derivs d;
            d.v = q0.x; d.dx = q0.y; d.dy = q1.x; d.ds = q1.y; d.dxx = q2.x;
            d.dyy = q2.y; d.dss = q3.x; d.dxy = q3.y; d.dxs = q4.x; d.dys = q4.y;
            for (int it = 0; it < MAX_INTERP; it++) {
                if (it > 0) d = derivs_at(J, o, l, (long long)r * C + c, C, RC);
                const float a = d.dxx, bb = d.dxy, cc = d.dxs, dd = d.dxy, e = d.dyy, ff = d.dys, gg = d.dxs,
                            h = d.dys, ii = d.dss;
                const float A0 = e * ii - ff * h, B0 = -(dd * ii - ff * gg), C0 = dd * h - e * gg;
                const float det = (a * A0 + bb * B0) + cc * C0;
                if (!(fabsf(det) > 1e-20f)) break;
                const float i00 = A0 / det, i01 = -(bb * ii - cc * h) / det, i02 = (bb * ff - cc * e) / det;
                const float i10 = B0 / det, i11 = (a * ii - cc * gg) / det, i12 = -(a * ff - cc * dd) / det;
                const float i20 = C0 / det, i21 = -(a * h - bb * gg) / det, i22 = (a * e - bb * dd) / det;
                const float x0 = -((i00 * d.dx + i01 * d.dy) + i02 * d.ds);
                const float x1 = -((i10 * d.dx + i11 * d.dy) + i12 * d.ds);
                const float x2 = -((i20 * d.dx + i21 * d.dy) + i22 * d.ds);
                if (fabsf(x0) < 0.5f && fabsf(x1) < 0.5f && fabsf(x2) < 0.5f) {
                    good = true;
                    X0 = x0; X1 = x1; X2 = x2;
                    break;
                }
                if (!(fabsf(x0) <= 1e6f && fabsf(x1) <= 1e6f && fabsf(x2) <= 1e6f)) break;   // also NaN
                c += (int)rintf(x0);
                r += (int)rintf(x1);
                l += (int)rintf(x2);
                if (l < 1 || l > N_LAYERS || c < BORDER || c >= C - BORDER || r < BORDER || r >= R - BORDER) break;
            }
            if (good) {
                // d is already at the final voxel: a converged step does not move
                const float contr = d.v + 0.5f * ((d.dx * X0 + d.dy * X1) + d.ds * X2);
                const float tr = d.dxx + d.dyy;
                const float dt = d.dxx * d.dyy - d.dxy * d.dxy;
                const float ac = fabsf(contr);
                if (ac * 3.0f >= J.contrast_f && dt > 0.f && (tr * tr) * J.edge_f < J.e1sq_f * dt) {
                    keep = true;
                    resp = ac;
                }
            }
        }
        int pos = wave_append(keep, &J.counters[1]);
        if (keep) {
            if (pos < J.cand_cap) {
                J.refined[2 * pos] = make_float4(__int_as_float(o), __int_as_float(l), __int_as_float(r), __int_as_float(c));
                J.refined[2 * pos + 1] = make_float4(X0, X1, X2, resp);
                J.refined_src[pos] = g;
            } else {
                atomicOr(&J.counters[3], 2);
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Gradient of a gauss layer at (y, x) (gsift.py:256-267): dx = G[y,x+1]-G[y,x-1],
// dy = G[y-1,x]-G[y+1,x]; magnitude sqrt (float32, correctly rounded); angle =
// float32(atan2) * (180f/pi_f) in [0, 360).  Callers only ask for interior pixels.

// Fixed-point (2^50 units) value of a non-negative contribution v < 2^2, rounded to nearest
// even: x = v * 2^50 < 2^52, so x + 2^52 lands in [2^52, 2^53) where the double's ulp is 1 and
// its mantissa bits are the rounded integer.  Same result as __double2ull_rn(v * 2^50) in two
// cheap operations (no float64 -> uint64 conversion sequence).
__device__ __forceinline__ unsigned long long fx50(double v) {
    const double y = v * 0x1p50 + 0x1p52;
    return (unsigned long long)__double_as_longlong(y) - 0x4330000000000000ull;
}

// Gradient angle in degrees [0, 360) for the descriptor's samples: t = min / max of |dx|, |dy|
// (one hardware reciprocal, 1 ulp), atan(t) = t * P(t^2) with an odd degree-17 polynomial
// fitted to atan on [0, 1] by iterated weighted least squares (atan_poly_fit.py in this
// directory; max error 1.0e-7 rad in float32 Horner form), then the octant folds.  1e-7 rad moves a sample by 5e-7 of an
// orientation bin; the descriptor's bar is +-1 unit per bin on the 512 scale.
__device__ __forceinline__ float desc_angle_deg(float dy, float dx) {
    const float ax = fabsf(dx), ay = fabsf(dy);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 1e-30f ? mn * __builtin_amdgcn_rcpf(mx) : 0.0f;
    const float z = t * t;
    float p = 0.00245657f;
    p = fmaf(p, z, -0.0144007f);
    p = fmaf(p, z, 0.03978004f);
    p = fmaf(p, z, -0.07234744f);
    p = fmaf(p, z, 0.10498885f);
    p = fmaf(p, z, -0.1416121f);
    p = fmaf(p, z, 0.19985904f);
    p = fmaf(p, z, -0.33332597f);
    p = fmaf(p, z, 0.99999989f);
    float r = t * p;   // (Horner steps fused)
    if (ay > ax) r = 1.57079632679489662f - r;
    if (dx < 0.0f) r = 3.14159265358979324f - r;
    const float d = r * (180.0f / 3.14159265358979324f);
    return dy < 0.0f ? 360.0f - d : d;
}

// fixed-point 2^-50 bits of a value already scaled by 2^50 (see fx50)
__device__ __forceinline__ unsigned long long fx50_scaled(double v) {
    return (unsigned long long)__double_as_longlong(v + 0x1p52) - 0x4330000000000000ull;
}

// Orientation (gsift.py:372-432 with the detect epilogue :462-469).  One wave per refined
// keypoint: 36-bin histogram of the gaussian-weighted gradient magnitudes in the
// (2*ORI_R_MAX+1)^2 window (own radius rint(4.5*scl)), smoothed, peaks >= 0.8 max with
// parabolic interpolation.  Each peak is appended as an oriented keypoint with its sort key.

constexpr float ORI_GUARD = 1e-4f;

// float32 rounding of the float64 atan2, kept out of line: inlined, its registers would be
// reserved for the whole orientation kernel although it runs for ~1 sample in 10^4
__device__ __attribute__((noinline)) float exact_atan2_f32(float dy, float dx) {
    return (float)atan2((double)dy, (double)dx);
}

// Oriented keypoints are staged in LDS and appended ORI_BUF at a time with one atomic on the
// image's counter (an atomic per keypoint, all on one address, cost more than the histograms).
constexpr int ORI_BUF = 64;
static_assert(ORI_BUF >= ORI_BINS, "one keypoint's peaks always fit an empty buffer");

__global__ __launch_bounds__(64) void k_orient(const SiftJob* __restrict__ jobs) {
    __shared__ unsigned long long s_h[ORI_BINS];
    __shared__ float4 s_rec[2 * ORI_BUF];
    __shared__ unsigned long long s_k0[ORI_BUF];
    __shared__ int s_rank[ORI_BUF];
    __shared__ float s_f[ORI_BINS];
    __shared__ float s_w[(ORI_R_MAX + 1) * (ORI_R_MAX + 2) / 2];   // gaussian weight by (max, min) of |i|, |j|
    const SiftJob& J = jobs[blockIdx.z];
    const int n = min(J.counters[1], J.cand_cap);
    const int lane = threadIdx.x;
    int nbuf = 0;   // staged keypoints (block = one wave: uniform)
    auto flush = [&]() {
        if (nbuf == 0) return;
        int base = 0;
        if (lane == 0) base = atomicAdd(&J.counters[2], nbuf);
        base = __shfl(base, 0);
        for (int i = lane; i < nbuf; i += 64) {
            const int pos = base + i;
            if (pos < J.kp_cap) {
                J.oriented[2 * pos] = s_rec[2 * i];
                J.oriented[2 * pos + 1] = s_rec[2 * i + 1];
                J.keys0[2 * pos] = s_k0[i];
                J.keys0[2 * pos + 1] = ((unsigned long long)s_rank[i] << 32) | (unsigned)pos;
            } else {
                atomicOr(&J.counters[3], 4);
            }
        }
        nbuf = 0;
        __syncthreads();   // the buffer is read before it is refilled
    };
    // the next keypoint's record is fetched while this one is processed
    float4 na = make_float4(0.f, 0.f, 0.f, 0.f), nb = na;
    if ((int)blockIdx.x < n) {
        na = J.refined[2 * blockIdx.x];
        nb = J.refined[2 * blockIdx.x + 1];
    }
    for (int k = blockIdx.x; k < n; k += gridDim.x) {
        const float4 a = na, b = nb;
        if (k + (int)gridDim.x < n) {
            na = J.refined[2 * (k + gridDim.x)];
            nb = J.refined[2 * (k + gridDim.x) + 1];
        }
        const int o = __float_as_int(a.x), l = __float_as_int(a.y), r = __float_as_int(a.z), c = __float_as_int(a.w);
        const int R = J.oct[o].R, C = J.oct[o].C;
        const float e = ((float)l + b.z) / 3.0f;
        const float scl = 1.6f * (float)pow(2.0, (double)e);
        const int rad = min((int)rintf(4.5f * scl), ORI_R_MAX);
        const float sw = 1.5f * scl;
        const float den = 2.0f * (sw * sw);
        const __attribute__((address_space(1))) float* G = gptr(J.G123 + J.oct[o].gbase + (long long)(l - 1) * R * C);
        if (lane < ORI_BINS) s_h[lane] = 0ull;
        // w = exp(-(i*i + j*j) / den), tabulated by (a, b) = (max(|i|, |j|), min(|i|, |j|)) at
        // a*(a+1)/2 + b: one float64 exp per pair that occurs, from the same d = a*a + b*b
        const int npair = (rad + 1) * (rad + 2) / 2;
        for (int t = lane; t < npair; t += 64) {
            int a = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
            while ((a + 1) * (a + 2) / 2 <= t) a++;
            while (a * (a + 1) / 2 > t) a--;
            const int b = t - a * (a + 1) / 2;
            s_w[t] = (float)exp((double)(-(float)(a * a + b * b) / den));
        }
        __syncthreads();
        // samples of the keypoint's own (2 rad + 1)^2 window (gsift.py:396-399 masks the rest)
        //
        // bin = rint(ANG * 0.1) with ANG from the float64-rounded atan2 (the reference's value).
        // atan2f is within 3 ulp of that rounding (fsx_atan2_probe, 2^28 pairs; a test asserts
        // <= 4): 4 ulp of |theta| <= pi is < 9.6e-7 rad; with the roundings of * deg, + 360 and
        // * 0.1, ANG * 0.1 then differs from the reference's by < 1.4e-5, so the fast bin is the
        // reference's unless ANG * 0.1 lies within ORI_GUARD = 1e-4 of a rounding boundary.
        // Such samples are only marked here (bit `it` of `late`; a lane has <= 18 samples) and
        // accumulated after the loop with the exact angle, which keeps the float64 atan2 and
        // its registers out of the hot loop.
        const int W = 2 * rad + 1;
        const float deg = 180.0f / 3.14159265358979323846f;
        auto sample = [&](int s, float& dx, float& dy, float& m) {
            const int i = s / W - rad, j = s - (s / W) * W - rad;
            const int y = r + i, x = c + j;
            if (!(y > 0 && y < R - 1 && x > 0 && x < C - 1)) return false;
            const size_t kk = (size_t)y * C + x;
            dx = G[kk + 1] - G[kk - 1];
            dy = G[kk - C] - G[kk + C];
            const float xx = dx * dx, yy = dy * dy;
            const int ai = abs(i), aj = abs(j), wa = max(ai, aj), wb = min(ai, aj);
            m = sqrtf(xx + yy) * s_w[wa * (wa + 1) / 2 + wb];
            return true;
        };
        auto accumulate = [&](float af, float m) {
            int bin = (int)rintf(af * 0.1f);
            bin = bin >= ORI_BINS ? bin - ORI_BINS : bin;
            atomicAdd(&s_h[bin], fx50((double)m));
        };
        static_assert((2 * ORI_R_MAX + 1) * (2 * ORI_R_MAX + 1) <= 64 * 32, "sample index fits the mask");
        unsigned late = 0u;
        // one iteration of software pipelining: the four gradient taps of the next sample are
        // loaded before this one is binned (the loads of an iteration otherwise wait alone)
        auto taps4 = [&](int s, float4& g, int& wi) {
            const int i = s / W - rad, j = s - (s / W) * W - rad;
            const int y = r + i, x = c + j;
            const bool ok = s < W * W && y > 0 && y < R - 1 && x > 0 && x < C - 1;
            if (ok) {
                const size_t kk = (size_t)y * C + x;
                g = make_float4(G[kk + 1], G[kk - 1], G[kk - C], G[kk + C]);
                const int ai = abs(i), aj = abs(j), wa = max(ai, aj), wb = min(ai, aj);
                wi = wa * (wa + 1) / 2 + wb;
            } else {
                wi = -1;
            }
        };
        float4 gn = make_float4(0.f, 0.f, 0.f, 0.f);
        int wn = -1;
        taps4(lane, gn, wn);
        for (int s = lane, it = 0; s < W * W; s += 64, ++it) {
            const float4 gc = gn;
            const int wc = wn;
            taps4(s + 64, gn, wn);
            if (wc < 0) continue;
            const float dx = gc.x - gc.y, dy = gc.z - gc.w;
            const float m = sqrtf(dx * dx + dy * dy) * s_w[wc];
            float af = atan2f(dy, dx) * deg;
            af = af < 0.f ? af + 360.0f : af;
            const float t = af * 0.1f;
            if (fabsf((t - floorf(t)) - 0.5f) < ORI_GUARD) {
                late |= 1u << it;
                continue;
            }
            accumulate(af, m);
        }
        while (late) {
            const int it = __builtin_ctz(late);
            late &= late - 1u;
            float dx, dy, m;
            sample(lane + 64 * it, dx, dy, m);
            const float ae = exact_atan2_f32(dy, dx) * deg;
            accumulate(ae < 0.f ? ae + 360.0f : ae, m);
        }
        __syncthreads();
        if (lane < ORI_BINS) s_f[lane] = (float)((double)s_h[lane] * 0x1p-50);
        __syncthreads();
        float h = -1.0f, hl = 0.f, hr = 0.f;
        if (lane < ORI_BINS) {
            auto H = [&](int j) { return s_f[(j + ORI_BINS) % ORI_BINS]; };
            h = ((H(lane - 2) + H(lane + 2)) * 0.0625f + (H(lane - 1) + H(lane + 1)) * 0.25f) + H(lane) * 0.375f;
        }
        __syncthreads();
        if (lane < ORI_BINS) s_f[lane] = h;
        __syncthreads();
        float hmax = h;
        for (int off = 32; off >= 1; off >>= 1) hmax = fmaxf(hmax, __shfl_xor(hmax, off));
        bool peak = false;
        float angle = 0.f;
        if (lane < ORI_BINS) {
            hl = s_f[(lane + ORI_BINS - 1) % ORI_BINS];
            hr = s_f[(lane + 1) % ORI_BINS];
            peak = h > hl && h > hr && h >= 0.8f * hmax && hmax > 0.0f;
            if (peak) {
                float dn = (hl - 2.0f * h) + hr;
                if (fabsf(dn) < 1e-12f) dn = -1e-12f;
                const float q = (0.5f * (hl - hr)) / dn;
                double bin = (double)lane + (double)q;
                if (bin < 0) bin += ORI_BINS;
                if (bin >= ORI_BINS) bin -= ORI_BINS;
                double ang = 360.0 - 10.0 * bin;
                if (fabs(ang - 360.0) < 1e-4) ang = 0.0;
                angle = (float)ang;
            }
        }
        const unsigned long long pm = __ballot(peak);
        const int rank = __popcll(pm & ((1ull << lane) - 1ull));
        const int npk = __popcll(pm);
        if (nbuf + npk > ORI_BUF) flush();
        if (peak) {
            const int i = nbuf + rank;
            const float f = (float)ldexp(1.0, o) * 0.5f;
            const float x = ((float)c + b.x) * f;
            const float y = ((float)r + b.y) * f;
            const float size = (scl * 2.0f) * f;
            const uint32_t src = J.refined_src[k];
            s_rec[2 * i] = make_float4(x, y, size, angle);
            s_rec[2 * i + 1] = make_float4(b.w, __int_as_float(o), __int_as_float(l), __uint_as_float(src));
            // stable argsort of -response (gsift.py:471): response descending, then
            // candidate order, then peak order within the keypoint (k1 gets the slot at flush)
            s_k0[i] = ((unsigned long long)(~__float_as_uint(b.w)) << 32) | src;
            s_rank[i] = rank;
        }
        nbuf += npk;
        __syncthreads();
    }
    flush();
}

// ------------------------------------------------------------------------------------------
// atan2f accuracy probe: the largest difference, in float32 ulps (distance along the ordered
// float32 line), between atan2f(y, x) and the float32 rounding of the float64 atan2 over n
// seeded argument pairs, plus the zero / equal-magnitude / sign cases.  Components are
// gradient-like: |y|, |x| < 1 with exponents 2^-1 .. 2^-40 (and, for 1 pair in 16, down to
// the float32 denormals).  The orientation kernel's fast bin path relies on this bound.

__device__ __forceinline__ unsigned long long probe_mix(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

__device__ __forceinline__ float probe_value(unsigned h, bool deep) {
    const unsigned sign = (h & 1u) << 31;
    const unsigned span = deep ? 150u : 40u;
    const int e = 126 - (int)((h >> 1) % span);                 // biased exponent 126 .. -23
    const unsigned mant = (h >> 9) & 0x7fffffu;
    const unsigned bits = e > 0 ? sign | ((unsigned)e << 23) | mant : sign | (mant >> (1 - e));
    return __uint_as_float(bits);
}

__device__ __forceinline__ int ordered_f32(float f) {
    const int a = __float_as_int(f);
    return a < 0 ? (int)(0x80000000u - (unsigned)a) : a;
}

__global__ __launch_bounds__(256) void k_atan2_probe(long long n, unsigned long long seed, unsigned* __restrict__ out) {
    unsigned m = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const unsigned long long h = probe_mix(seed ^ probe_mix((unsigned long long)i));
        const bool deep = (h >> 60) == 0;
        float y = probe_value((unsigned)h, deep), x = probe_value((unsigned)(h >> 32), deep);
        switch (i & 63) {                                      // special cases
            case 0: x = 0.0f; break;
            case 1: y = 0.0f; break;
            case 2: y = x; break;
            case 3: y = -x; break;
            case 4: x = -0.0f; break;
            default: break;
        }
        const float f = atan2f(y, x);
        const float r = (float)atan2((double)y, (double)x);
        const int d = ordered_f32(f) - ordered_f32(r);
        m = max(m, (unsigned)(d < 0 ? -d : d));
    }
    for (int off = 32; off >= 1; off >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, off));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// The descriptor's angle (desc_angle_deg) against the float64 atan2 in degrees on [0, 360):
// out[0] = largest circular difference in units of 1e-9 degree over the same argument pairs
// (only normal-range components: the descriptor's gradients are differences of Gaussian
// values in [0, 1]), out[1] = the same for the atan2f form, for comparison.
__global__ __launch_bounds__(256) void k_desc_angle_probe(long long n, unsigned long long seed,
                                                          unsigned* __restrict__ out) {
    unsigned m0 = 0, m1 = 0;
    const float deg = 180.0f / 3.14159265358979323846f;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
        const unsigned long long h = probe_mix(seed ^ probe_mix((unsigned long long)i));
        float y = probe_value((unsigned)h, false), x = probe_value((unsigned)(h >> 32), false);
        switch (i & 63) {
            case 0: x = 0.0f; break;
            case 1: y = 0.0f; break;
            case 2: y = x; break;
            case 3: y = -x; break;
            default: break;
        }
        double r = atan2((double)y, (double)x) * (180.0 / 3.14159265358979323846);
        if (r < 0.0) r += 360.0;
        const float a0 = desc_angle_deg(y, x);
        float a1 = atan2f(y, x) * deg;
        a1 = a1 < 0.0f ? a1 + 360.0f : a1;
        double d0 = fabs((double)a0 - r), d1 = fabs((double)a1 - r);
        d0 = fmin(d0, 360.0 - d0);
        d1 = fmin(d1, 360.0 - d1);
        m0 = max(m0, (unsigned)ceil(d0 * 1e9));
        m1 = max(m1, (unsigned)ceil(d1 * 1e9));
    }
    for (int off = 32; off >= 1; off >>= 1) {
        m0 = max(m0, (unsigned)__shfl_xor((int)m0, off));
        m1 = max(m1, (unsigned)__shfl_xor((int)m1, off));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&out[0], m0);
        atomicMax(&out[1], m1);
    }
}

// ------------------------------------------------------------------------------------------
// Deterministic sort of (k0, k1) key pairs: 1024-element bitonic blocks in LDS, then merge
// passes where every element finds its output slot by binary search in the partner run.

__device__ __forceinline__ bool key_less(unsigned long long a0, unsigned long long a1, unsigned long long b0,
                                         unsigned long long b1) {
    return a0 < b0 || (a0 == b0 && a1 < b1);
}

__global__ __launch_bounds__(256) void k_sort_local(const SiftJob* __restrict__ jobs) {
    __shared__ unsigned long long s0[1024], s1[1024];
    const SiftJob& J = jobs[blockIdx.z];
    const int n = min(J.counters[2], J.kp_cap);
    const int base = blockIdx.x * 1024;
    if (base >= n) return;
    for (int t = threadIdx.x; t < 1024; t += 256) {
        int g = base + t;
        if (g < n) {
            s0[t] = J.keys0[2 * g];
            s1[t] = J.keys0[2 * g + 1];
        } else {
            s0[t] = ~0ull;
            s1[t] = ~0ull;
        }
    }
    __syncthreads();
    for (int k = 2; k <= 1024; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int t = threadIdx.x; t < 512; t += 256) {
                int i = 2 * t - (t & (j - 1));
                int p = i + j;
                bool up = (i & k) == 0;
                unsigned long long a0 = s0[i], a1 = s1[i], b0 = s0[p], b1 = s1[p];
                bool sw = up ? key_less(b0, b1, a0, a1) : key_less(a0, a1, b0, b1);
                if (sw) {
                    s0[i] = b0; s1[i] = b1;
                    s0[p] = a0; s1[p] = a1;
                }
            }
            __syncthreads();
        }
    }
    for (int t = threadIdx.x; t < 1024; t += 256) {
        int g = base + t;
        if (g < n) {
            J.keys0[2 * g] = s0[t];
            J.keys0[2 * g + 1] = s1[t];
        }
    }
}

__global__ __launch_bounds__(256) void k_merge(const SiftJob* __restrict__ jobs, int w, int src_is_0) {
    const SiftJob& J = jobs[blockIdx.z];
    const int n = min(J.counters[2], J.kp_cap);
    const unsigned long long* src = src_is_0 ? J.keys0 : J.keys1;
    unsigned long long* dst = src_is_0 ? J.keys1 : J.keys0;
    for (int p = blockIdx.x * 256 + threadIdx.x; p < n; p += gridDim.x * 256) {
        const unsigned long long x0 = src[2 * p], x1 = src[2 * p + 1];
        const int s = (p / (2 * w)) * (2 * w);
        int lo, hi, out;
        if (p < s + w) {
            lo = s + w;
            hi = min(s + 2 * w, n);
            out = p;
        } else {
            lo = s;
            hi = s + w;
            out = p - w;
        }
        // stable merge: a left-run element counts right-run keys < x, a right-run element
        // counts left-run keys <= x (equal keys keep their order and never collide)
        const bool left = p < s + w;
        int a = lo, b = hi;
        while (a < b) {
            int m = (a + b) >> 1;
            const unsigned long long m0 = src[2 * m], m1 = src[2 * m + 1];
            const bool before = left ? key_less(m0, m1, x0, x1) : !key_less(x0, x1, m0, m1);
            if (before) a = m + 1;
            else b = m;
        }
        out += a - lo;
        dst[2 * out] = x0;
        dst[2 * out + 1] = x1;
    }
}

// emit the first min(n, max_kp) keypoints in sorted order as the detect() arrays
__global__ __launch_bounds__(256) void k_emit(const SiftJob* __restrict__ jobs, int final_in_0) {
    const SiftJob& J = jobs[blockIdx.z];
    const int n = min(J.counters[2], J.kp_cap);
    const int m = min(n, J.max_kp);
    const unsigned long long* keys = final_in_0 ? J.keys0 : J.keys1;
    const int stride = J.max_kp;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < m; i += gridDim.x * 256) {
        const int idx = (int)(keys[2 * i + 1] & 0xffffffffull);
        const float4 a = J.oriented[2 * idx], b = J.oriented[2 * idx + 1];
        float* out = J.kp_out;
        out[i] = a.x;
        out[stride + i] = a.y;
        out[2 * stride + i] = a.z;
        out[3 * stride + i] = a.w;
        out[4 * stride + i] = b.x;
        ((int*)out)[5 * stride + i] = __float_as_int(b.y);
        ((int*)out)[6 * stride + i] = __float_as_int(b.z);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) *J.n_out = m;
}

__global__ void k_reset(const SiftJob* __restrict__ jobs) {
    const SiftJob& J = jobs[blockIdx.x];
    if (threadIdx.x < 8) J.counters[threadIdx.x] = 0;
}

// ------------------------------------------------------------------------------------------
// Descriptor (gsift.py:480-525 describe, 800-871 _descr_flat).  One 256-thread block per
// selected keypoint.  Sample geometry in float64 (the reference's host-side scale factors
// are float64), angle bin in float32, trilinear split into the (6,6,10) histogram with
// 64-bit fixed-point LDS atomics, then the orientation wrap, the 4x4x8 crop and numpy's
// float32 normalisation (pairwise sums with 8 accumulators): L2 -> clip 0.2 -> L2 -> x512.
// Optional RootSIFT uint8 epilogue (features.py:80-82).

struct DescArgs {
    const float* kp;      // [7][stride] detect arrays (x,y,size,angle,response,octave,layer)
    int stride;
    const long long* sel; // selected keypoint indices
    int m;
    float* desc;          // (m, 128)
    uint8_t* u8;          // (m, 128) or null
};

__device__ __forceinline__ float wave_pairwise8(float v_lane_j_partial) {
    // lanes 0..7 hold numpy's 8 running accumulators; combine in numpy's fixed tree
    float r0 = __shfl(v_lane_j_partial, 0), r1 = __shfl(v_lane_j_partial, 1);
    float r2 = __shfl(v_lane_j_partial, 2), r3 = __shfl(v_lane_j_partial, 3);
    float r4 = __shfl(v_lane_j_partial, 4), r5 = __shfl(v_lane_j_partial, 5);
    float r6 = __shfl(v_lane_j_partial, 6), r7 = __shfl(v_lane_j_partial, 7);
    return ((r0 + r1) + (r2 + r3)) + ((r4 + r5) + (r6 + r7));
}

constexpr int DS_COPIES = 4;   // interleaved copies of the descriptor histogram (a sweep of 1..16: 4 is fastest)

// One keypoint's descriptor in pieces: the keypoint's frame (DescKP), the column span of a
// window row (desc_span), one sample's gradient and its eight trilinear votes (desc_sample),
// and the histogram to descriptor step (desc_finish, by one wave).  The votes are fixed-point
// integers, so the histogram does not depend on which thread adds which sample.  (One wave
// per keypoint instead of a 256-thread block was measured slower: 1.1-1.25 ms against 0.7 ms
// for the 8-photo step, with 1, 2 or 4 histogram copies per wave.)
struct DescKP {
    bool ok;                        // false: a selection outside the arrays, or an octave / layer
                                    // the workspace does not hold (a zero row)
    int R, C, pr32, pc32;
    long long rad;
    float ori;
    double cos_t, sin_t;
    const __attribute__((address_space(1))) float* G;
};

__device__ __forceinline__ DescKP desc_frame(const SiftJob& J, const float* __restrict__ kpa_, int stride,
                                             long long k) {
    const __attribute__((address_space(1))) float* kpa = gptr(kpa_);
    const __attribute__((address_space(1))) int* kpi = gptr((const int*)kpa_);
    DescKP D;
    const bool k_ok = k >= 0 && k < stride;
    const int o = k_ok ? kpi[5 * (long long)stride + k] : -1, li = k_ok ? kpi[6 * (long long)stride + k] : 0;
    D.ok = o >= 0 && o < J.n_oct && li >= 1 && li <= N_LAYERS;
    if (!D.ok) return D;
    const int R = J.oct[o].R, C = J.oct[o].C;
    const double f = 1.0 / (ldexp(1.0, o) * 0.5);
    const double px = (double)kpa[k] * f, py = (double)kpa[(long long)stride + k] * f;
    const double scl = (double)kpa[2 * (long long)stride + k] * f * 0.5;
    float ori = 360.0f - kpa[3 * (long long)stride + k];
    if (fabsf(ori - 360.0f) < 1e-4f) ori = 0.0f;
    long long rad = (long long)rint(((3.0 * scl) * 1.4142135623730951) * 5 * 0.5);
    const long long rlim = (long long)sqrt((double)((long long)R * R + (long long)C * C));
    rad = rad < rlim ? rad : rlim;
    rad = rad < 23000 ? rad : 23000;   // (2 rad + 1)^2 stays below 2^31
    const double hw = 3.0 * scl;
    const float t = ori * (float)(3.14159265358979323846 / 180.0);
    D.cos_t = (double)(float)cos((double)t) / hw;
    D.sin_t = (double)(float)sin((double)t) / hw;
    D.pr32 = (int)(long long)rint(py);     // keypoints lie inside the octave
    D.pc32 = (int)(long long)rint(px);
    D.R = R;
    D.C = C;
    D.rad = rad;
    D.ori = ori;
    D.G = gptr(J.G123 + J.oct[o].gbase + (long long)(li - 1) * R * C);
    return D;
}

// window row `row`: the column span [jl, jl + n) outside of which |c_rot| or |r_rot| >= 2.5 for
// sure (widened by one sample; the exact per-sample test still decides), so only the rotated
// 5x5-cell square is visited instead of the whole (2 rad + 1)^2 window
__device__ __forceinline__ void desc_span(const DescKP& D, int row, int& jl, int& n) {
    const double ii = (double)(row - D.rad);
    double lo = -(double)D.rad, hi = (double)D.rad;
    // c_rot = j cos_t - ii sin_t in (-2.5, 2.5); r_rot = j sin_t + ii cos_t in (-2.5, 2.5)
    const double cc[2] = {D.cos_t, D.sin_t}, off[2] = {-ii * D.sin_t, ii * D.cos_t};
    for (int q = 0; q < 2; q++) {
        if (cc[q] != 0.0) {
            const double a = (-2.5 - off[q]) / cc[q], b2 = (2.5 - off[q]) / cc[q];
            lo = fmax(lo, floor(fmin(a, b2)) - 1.0);
            hi = fmin(hi, ceil(fmax(a, b2)) + 1.0);
        } else if (!(fabs(off[q]) < 2.5)) {
            hi = lo - 1.0;
        }
    }
    jl = (int)lo;
    n = hi >= lo ? (int)(hi - lo) + 1 : 0;
}

// sample (i, j) of the window (offsets from the keypoint): gradient as gsift.py:256-267
// (magnitude correctly rounded, angle by desc_angle_deg), Gaussian weight, and the eight
// trilinear votes into copy `cp` of the NC interleaved copies of the histogram s_h
template <int NC>
__device__ __forceinline__ void desc_sample(const DescKP& D, int i, int j, unsigned long long* s_h, int cp) {
    const float obs = (float)(8.0 / 360.0);
    const double oi = (double)(float)i, oj = (double)(float)j;
    const double c_rot = oj * D.cos_t - oi * D.sin_t;
    const double r_rot = oj * D.sin_t + oi * D.cos_t;
    const double rbin = (r_rot + 2.0) - 0.5, cbin = (c_rot + 2.0) - 0.5;
    const int yy = D.pr32 + i, xx = D.pc32 + j;
    if (!(rbin > -1.0 && rbin < 4.0 && cbin > -1.0 && cbin < 4.0 && yy > 0 && yy < D.R - 1 && xx > 0 && xx < D.C - 1))
        return;
    const size_t kk = (size_t)yy * D.C + xx;
    const float gdx = D.G[kk + 1] - D.G[kk - 1];
    const float gdy = D.G[kk - D.C] - D.G[kk + D.C];
    const float mag = sqrtf(gdx * gdx + gdy * gdy);
    const float ang = desc_angle_deg(gdy, gdx);
    const double w = (double)__expf((float)((c_rot * c_rot + r_rot * r_rot) * (-0.125)));
    const double mg = ((double)mag * w) * 0x1p50;   // exact scaling: fixed point below
    const float obin = (ang - D.ori) * obs;
    const double r0 = floor(rbin), c0 = floor(cbin);
    const float o0f = floorf(obin);
    const double rb = rbin - r0, cb = cbin - c0;
    const double ob = (double)(obin - o0f);
    int o0 = (int)o0f;
    if (o0 < 0) o0 += 8;
    if (o0 >= 8) o0 -= 8;
    const double v_r1 = mg * rb, v_r0 = mg - v_r1;
    const double v_rc11 = v_r1 * cb, v_rc10 = v_r1 - v_rc11;
    const double v_rc01 = v_r0 * cb, v_rc00 = v_r0 - v_rc01;
    const double v111 = v_rc11 * ob, v110 = v_rc11 - v111;
    const double v101 = v_rc10 * ob, v100 = v_rc10 - v101;
    const double v011 = v_rc01 * ob, v010 = v_rc01 - v011;
    const double v001 = v_rc00 * ob, v000 = v_rc00 - v001;
    // neighbouring samples, which mostly share their bins, spread over NC addresses instead of
    // contending for one
    const int idx = ((((int)r0 + 1) * 6 + (int)c0 + 1) * 10 + o0) * NC + cp;
    atomicAdd(&s_h[idx], fx50_scaled(v000));
    atomicAdd(&s_h[idx + 1 * NC], fx50_scaled(v001));
    atomicAdd(&s_h[idx + 10 * NC], fx50_scaled(v010));
    atomicAdd(&s_h[idx + 11 * NC], fx50_scaled(v011));
    atomicAdd(&s_h[idx + 60 * NC], fx50_scaled(v100));
    atomicAdd(&s_h[idx + 61 * NC], fx50_scaled(v101));
    atomicAdd(&s_h[idx + 70 * NC], fx50_scaled(v110));
    atomicAdd(&s_h[idx + 71 * NC], fx50_scaled(v111));
}

// descriptor bin b (0..127) from the NC copies: exact integer sum, the orientation wrap bins
// folded in (bins 8, 9 of a cell are orientations 0, 1 again)
template <int NC>
__device__ __forceinline__ float desc_bin(const unsigned long long* s_h, int b) {
    const int rr = b >> 5, cc = (b >> 3) & 3, ob = b & 7;
    const int cell = ((rr + 1) * 6 + cc + 1) * 10;
    unsigned long long h0 = 0ull, h1 = 0ull;   // exact integer sums of the copies
#pragma unroll
    for (int q = 0; q < NC; q++) {
        h0 += s_h[(cell + ob) * NC + q];
        h1 += s_h[(cell + 8 + ob) * NC + q];
    }
    float v = (float)((double)h0 * 0x1p-50);
    if (ob < 2) v = v + (float)((double)h1 * 0x1p-50);
    return v;
}

__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// bins d0 = bin lane, d1 = bin lane + 64 of a descriptor -> its rows (one wave; s_d: 128 floats
// of LDS): clip at 0.2 of the L2 norm, renormalise to 512 (features.py / gsift.py), then the
// RootSIFT uint8 encoding when u8_row is set
__device__ __forceinline__ void desc_finish(float d0, float d1, float* s_d, float* __restrict__ desc_row,
                                            uint8_t* __restrict__ u8_row) {
    const int lane = threadIdx.x & 63;
    wave_lds_sync();
    s_d[lane] = d0;
    s_d[lane + 64] = d1;
    wave_lds_sync();
    float part = 0.f;
    if (lane < 8) {
        part = s_d[lane] * s_d[lane];
        for (int t2 = 1; t2 < 16; t2++) {
            const float v = s_d[8 * t2 + lane];
            part += v * v;
        }
    }
    const float nrm = sqrtf(wave_pairwise8(part));
    const float thr = nrm * 0.2f;
    d0 = d0 < thr ? d0 : thr;
    d1 = d1 < thr ? d1 : thr;
    wave_lds_sync();
    s_d[lane] = d0;
    s_d[lane + 64] = d1;
    wave_lds_sync();
    part = 0.f;
    if (lane < 8) {
        part = s_d[lane] * s_d[lane];
        for (int t2 = 1; t2 < 16; t2++) {
            const float v = s_d[8 * t2 + lane];
            part += v * v;
        }
    }
    const float nrm2 = sqrtf(wave_pairwise8(part));
    const float sc = 512.0f / (nrm2 > 1e-7f ? nrm2 : 1e-7f);
    const float e0 = d0 * sc, e1 = d1 * sc;
    if (desc_row) {
        desc_row[lane] = e0;
        desc_row[lane + 64] = e1;
    }
    if (u8_row) {
        wave_lds_sync();
        s_d[lane] = e0;
        s_d[lane + 64] = e1;
        wave_lds_sync();
        part = 0.f;
        if (lane < 8) {
            part = fabsf(s_d[lane]);
            for (int t2 = 1; t2 < 16; t2++) part += fabsf(s_d[8 * t2 + lane]);
        }
        const float l1 = wave_pairwise8(part) + 1e-7f;
        float u0 = sqrtf(e0 / l1) * 512.0f, u1 = sqrtf(e1 / l1) * 512.0f;
        u0 = u0 < 0.f ? 0.f : (u0 > 255.f ? 255.f : u0);
        u1 = u1 < 0.f ? 0.f : (u1 > 255.f ? 255.f : u1);
        u8_row[lane] = (uint8_t)u0;
        u8_row[lane + 64] = (uint8_t)u1;
    }
    wave_lds_sync();
}

__device__ __forceinline__ void desc_zero_row(float* __restrict__ desc_row, uint8_t* __restrict__ u8_row, int t) {
    if (t < 128) {
        if (desc_row) desc_row[t] = 0.0f;
        if (u8_row) u8_row[t] = 0;
    }
}

// one keypoint k of the arrays kp[7][stride], by a 256-thread block; s_h: HIST_SIZE * DS_COPIES
// u64, s_d: 128 floats of LDS
__device__ __forceinline__ void describe_kp(const SiftJob& J, const float* __restrict__ kpa_, int stride, long long k,
                                           float* __restrict__ desc_row, uint8_t* __restrict__ u8_row,
                                           unsigned long long* s_h, float* s_d) {
    const DescKP D = desc_frame(J, kpa_, stride, k);     // block-uniform
    if (!D.ok) {
        desc_zero_row(desc_row, u8_row, threadIdx.x);
        return;
    }
    const long long rad = D.rad;
    for (int b = threadIdx.x; b < HIST_SIZE * DS_COPIES; b += 256) s_h[b] = 0ull;
    __shared__ int s_jl[2 * 38 + 1 + 2], s_pre[2 * 38 + 2 + 2], s_wtot;
    __shared__ uint8_t s_row[(2 * 38 + 1) * (2 * 38 + 1)];   // window row of each span sample
    const int nrow = (int)(2 * rad + 1);
    const bool spans = nrow <= 2 * 38 + 1;
    if (spans) {
        for (int row = threadIdx.x; row < nrow; row += 256) {
            int jl, n;
            desc_span(D, row, jl, n);
            s_jl[row] = jl;
            s_pre[row + 1] = n;
        }
        __syncthreads();
        // inclusive scan of the per-row counts by the first two waves (nrow <= 77 < 128)
        int v = 0;
        if (threadIdx.x < 128) {
            v = (int)threadIdx.x < nrow ? s_pre[threadIdx.x + 1] : 0;
            const int ln = threadIdx.x & 63;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(v, off);
                if (ln >= off) v += u;
            }
            if (threadIdx.x == 63) s_wtot = v;
        }
        __syncthreads();
        if (threadIdx.x < 128) {
            if (threadIdx.x >= 64) v += s_wtot;
            if ((int)threadIdx.x < nrow) s_pre[threadIdx.x + 1] = v;
        }
        if (threadIdx.x == 0) s_pre[0] = 0;
        __syncthreads();
        // the window row of every span sample (replaces a binary search per sample)
        for (int row = threadIdx.x; row < nrow; row += 256)
            for (int q = s_pre[row]; q < s_pre[row + 1]; q++) s_row[q] = (uint8_t)row;
    }
    __syncthreads();
    const int W = (int)(2 * rad + 1);
    const int total = spans ? s_pre[nrow] : W * W;
    for (int s = threadIdx.x; s < total; s += 256) {
        int i, j;
        if (spans) {
            const int lo = s_row[s];
            i = lo - (int)rad;
            j = s_jl[lo] + (s - s_pre[lo]);
        } else {
            i = s / W - (int)rad;
            j = s - (s / W) * W - (int)rad;
        }
        desc_sample<DS_COPIES>(D, i, j, s_h, threadIdx.x % DS_COPIES);
    }
    __syncthreads();
    if (threadIdx.x < 128) s_d[threadIdx.x] = desc_bin<DS_COPIES>(s_h, threadIdx.x);
    __syncthreads();
    if (threadIdx.x < 64) desc_finish(s_d[threadIdx.x], s_d[threadIdx.x + 64], s_d, desc_row, u8_row);
    __syncthreads();
}

__global__ __launch_bounds__(256) void k_describe(const SiftJob* __restrict__ job, DescArgs A) {
    __shared__ unsigned long long s_h[HIST_SIZE * DS_COPIES];
    __shared__ float s_d[128];
    const int q = blockIdx.x;
    if (q >= A.m) return;
    describe_kp(*job, A.kp, A.stride, A.sel[q], A.desc + (long long)q * 128, A.u8 ? A.u8 + (long long)q * 128 : nullptr,
                s_h, s_d);
}

// feature batch: descriptor q of image blockIdx.z for its spread selection, in final order
__global__ __launch_bounds__(256) void k_describe_sel(const SiftJob* __restrict__ jobs) {
    __shared__ unsigned long long s_h[HIST_SIZE * DS_COPIES];
    __shared__ float s_d[128];
    const SiftJob& J = jobs[blockIdx.z];
    const int q = blockIdx.x;
    if (q >= J.counters[4]) return;
    describe_kp(J, J.kp_out, J.max_kp, J.sel[q], J.desc_out ? J.desc_out + (long long)q * 128 : nullptr,
                J.u8_out ? J.u8_out + (long long)q * 128 : nullptr, s_h, s_d);
}

// ------------------------------------------------------------------------------------------
// Grid spread of the detected keypoints (features.py:38-57 _spread, then the response sort
// of features.py:126).  One 256-thread block per image.  With N detected (already sorted by
// response, ties by detection order) and n wanted: cell = 12x9 grid cell of (x, y); the
// keypoints of a cell are ranked in detection order; ranks < ceil(n/108) are kept in (cell,
// rank) order, then the best of the rest fill up to n; finally a stable sort by -response.

constexpr int SP_THREADS = 1024;
constexpr int SP_SEGS = 256;
constexpr int SP_LDS_N = 32768;

__device__ int block_exclusive_scan(int v, int* s_tmp, int& total) {
    // SP_THREADS threads; s_tmp: SP_THREADS / 64 ints
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int x = v;
    for (int off = 1; off < 64; off <<= 1) {
        int y = __shfl_up(x, off);
        if (lane >= off) x += y;
    }
    if (lane == 63) s_tmp[wv] = x;
    __syncthreads();
    int wsum = 0;
    total = 0;
    for (int w2 = 0; w2 < SP_THREADS / 64; w2++) {
        if (w2 < wv) wsum += s_tmp[w2];
        total += s_tmp[w2];
    }
    __syncthreads();
    return wsum + x - v;
}

__global__ __launch_bounds__(SP_THREADS) void k_spread(const SiftJob* __restrict__ jobs) {
    __shared__ int s_cnt[108];
    __shared__ int s_start[109];
    __shared__ int s_tmp[SP_THREADS / 64];
    __shared__ unsigned short s_seg[SP_SEGS][108];
    __shared__ uint8_t s_cell[SP_LDS_N];   // cell ids for the serial segment passes (N <= SP_LDS_N)
    const SiftJob& J = jobs[blockIdx.z];
    const int N = *J.n_out, n = J.nfeat, stride = J.max_kp;
    const int tid = threadIdx.x;
    if (N <= n) {
        for (int i = tid; i < N; i += SP_THREADS) J.sel[i] = i;
        if (tid == 0) J.counters[4] = N;
        return;
    }
    const float* x = J.kp_out;
    const float* y = J.kp_out + stride;
    const float* resp = J.kp_out + 4 * (long long)stride;
    int* cell = J.spread_tmp;
    int* rank = cell + J.kp_cap;
    int* spos = rank + J.kp_cap;
    int* sbefore = spos + J.kp_cap;
    const float Wf = (float)J.W, Hf = (float)J.H;
    const bool lds = N <= SP_LDS_N;
    for (int i = tid; i < N; i += SP_THREADS) {
        int cx = (int)((x[i] / Wf) * 12.0f), cy = (int)((y[i] / Hf) * 9.0f);
        cx = cx < 0 ? 0 : (cx > 11 ? 11 : cx);
        cy = cy < 0 ? 0 : (cy > 8 ? 8 : cy);
        cell[i] = cy * 12 + cx;
        if (lds) s_cell[i] = (uint8_t)(cy * 12 + cx);
    }
    __syncthreads();
    // rank of each keypoint within its cell in detection order: SP_SEGS threads each own a
    // contiguous segment (< 65536 keypoints per segment), count per cell, exclusive-scan the
    // counts over the segments, then rank their own segment
    const int seg = (N + SP_SEGS - 1) / SP_SEGS;
    if (tid < SP_SEGS) {
        for (int c = 0; c < 108; c++) s_seg[tid][c] = 0;
        const int a = tid * seg, b = min(N, a + seg);
        if (lds)
            for (int i = a; i < b; i++) s_seg[tid][s_cell[i]]++;
        else
            for (int i = a; i < b; i++) s_seg[tid][cell[i]]++;
    }
    __syncthreads();
    if (tid < 108) {
        int run = 0;
        for (int t = 0; t < SP_SEGS; t++) {
            const int v = s_seg[t][tid];
            s_seg[t][tid] = (unsigned short)run;   // < 65536: the host caps max_kp
            run += v;
        }
        s_cnt[tid] = run;
    }
    __syncthreads();
    if (tid < SP_SEGS) {
        const int a = tid * seg, b = min(N, a + seg);
        if (lds)
            for (int i = a; i < b; i++) rank[i] = s_seg[tid][s_cell[i]]++;
        else
            for (int i = a; i < b; i++) rank[i] = s_seg[tid][cell[i]]++;
    }
    __syncthreads();
    const int cap = (n + 107) / 108;
    if (tid == 0) {
        s_start[0] = 0;
        for (int c = 0; c < 108; c++) s_start[c + 1] = s_start[c] + min(s_cnt[c], cap);
    }
    __syncthreads();
    const int K = s_start[108];
    // positions in the spread list: kept ones by (cell, rank), the rest after them in
    // detection order; selected = position < n
    int carry = 0;
    for (int base = 0; base < N; base += SP_THREADS) {
        const int i = base + tid;
        const bool kept = i < N && rank[i] < cap;
        int tot;
        const int ex = block_exclusive_scan((i < N && !kept) ? 1 : 0, s_tmp, tot);
        if (i < N) spos[i] = kept ? s_start[cell[i]] + rank[i] : K + carry + ex;
        carry += tot;
    }
    __syncthreads();
    carry = 0;
    for (int base = 0; base < N; base += SP_THREADS) {
        const int i = base + tid;
        const bool selq = i < N && spos[i] < n;
        int tot;
        const int ex = block_exclusive_scan(selq ? 1 : 0, s_tmp, tot);
        if (i < N) sbefore[i] = carry + ex;
        carry += tot;
    }
    __syncthreads();
    for (int i = tid; i < N; i += SP_THREADS) {
        const int p = spos[i];
        if (p >= n) continue;
        const float ri = resp[i];
        int s0 = i, e0 = i + 1;
        while (s0 > 0 && resp[s0 - 1] == ri) s0--;
        while (e0 < N && resp[e0] == ri) e0++;
        int r = 0;
        for (int j = s0; j < e0; j++)
            if (spos[j] < n && spos[j] < p) r++;
        J.sel[sbefore[s0] + r] = i;
    }
    if (tid == 0) J.counters[4] = n;
}

// ------------------------------------------------------------------------------------------
// Host side

static inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

struct WsLayout {
    size_t job, counters, G0, G123, G45, cand, cand_der, refined, refined_src, oriented, keys0, keys1, sel, spread, total;
};

static WsLayout ws_layout(const SiftPlan& p, int cand_cap, int kp_cap) {
    WsLayout L;
    size_t off = 0;
    L.job = off; off = align256(off + sizeof(SiftJob));
    L.counters = off; off = align256(off + 16 * sizeof(int));
    L.G0 = off; off = align256(off + (size_t)p.sumP * 4);
    L.G123 = off; off = align256(off + 3 * (size_t)p.sumP * 4);
    L.G45 = off; off = align256(off + 2 * (size_t)p.sumP * 4);
    L.cand = off; off = align256(off + (size_t)cand_cap * 4);
    L.cand_der = off; off = align256(off + (size_t)cand_cap * DER_N * 4);
    L.refined = off; off = align256(off + (size_t)cand_cap * 32);
    L.refined_src = off; off = align256(off + (size_t)cand_cap * 4);
    L.oriented = off; off = align256(off + (size_t)kp_cap * 32);
    L.keys0 = off; off = align256(off + (size_t)kp_cap * 16);
    L.keys1 = off; off = align256(off + (size_t)kp_cap * 16);
    L.sel = off; off = align256(off + (size_t)kp_cap * 8);
    L.spread = off; off = align256(off + (size_t)kp_cap * 16);
    L.total = off;
    return L;
}

static int fill_job(SiftJob* J, const SiftPlan& p, const uint8_t* img, int ld, double contrast, double edge,
                    int max_kp, int cand_cap, int kp_cap, void* ws, float* kp_out, int* n_out) {
    WsLayout L = ws_layout(p, cand_cap, kp_cap);
    char* b = (char*)ws;
    memset(J, 0, sizeof(*J));
    J->img = img;
    J->H = p.H;
    J->W = p.W;
    J->ld = ld;
    J->n_oct = p.n_oct;
    long long g45 = 0, gb = 0, g0 = 0;
    for (int o = 0; o < p.n_oct; o++) {
        long long P = (long long)p.R[o] * p.C[o];
        J->oct[o].R = p.R[o];
        J->oct[o].C = p.C[o];
        J->oct[o].g45base = g45;
        J->oct[o].gbase = gb;
        J->oct[o].g0base = g0;
        g45 += 2 * P;
        gb += 3 * P;
        g0 += P;
    }
    J->G0 = (float*)(b + L.G0);
    J->G123 = (float*)(b + L.G123);
    J->G45 = (float*)(b + L.G45);
    J->cand = (uint32_t*)(b + L.cand);
    J->cand_der = (float*)(b + L.cand_der);
    J->refined = (float4*)(b + L.refined);
    J->refined_src = (uint32_t*)(b + L.refined_src);
    J->oriented = (float4*)(b + L.oriented);
    J->keys0 = (unsigned long long*)(b + L.keys0);
    J->keys1 = (unsigned long long*)(b + L.keys1);
    J->counters = (int*)(b + L.counters);
    J->sel = (long long*)(b + L.sel);
    J->spread_tmp = (int*)(b + L.spread);
    J->cand_cap = cand_cap;
    J->kp_cap = kp_cap;
    // gsift.py:295 prefilter threshold is float64 then compared in float32; contrast and
    // edge are weak Python scalars (float32) in the tests of gsift.py:367
    J->pthr = (float)(0.8 * 0.5 * contrast / N_LAYERS);
    J->contrast_f = (float)contrast;
    J->edge_f = (float)edge;
    J->e1sq_f = (float)((edge + 1.0) * (edge + 1.0));
    J->max_kp = max_kp;
    J->kp_out = kp_out;
    J->n_out = n_out;
    return FSX_OK;
}

// one blur launch with the tuned row strip (TUNE_BLUR_ROWS: 8, 16, 24 or 32 output rows per tile)
template <int R2, bool BASE>
static void launch_blur(const SiftJob* jobs, int n_img, int o, int layer, int R, int C, const HalfTaps& h,
                        hipStream_t st) {
    constexpr int TWo = blur_tw<R2>();
    const int th = tune_get(TUNE_BLUR_ROWS);
    dim3 grid((C + TWo - 1) / TWo, (R + th - 1) / th, n_img);
    if (th == 8) hipLaunchKernelGGL((k_blur<R2, BASE, 8>), grid, dim3(BL_NT), 0, st, jobs, o, layer, h);
    else if (th == 24) hipLaunchKernelGGL((k_blur<R2, BASE, 24>), grid, dim3(BL_NT), 0, st, jobs, o, layer, h);
    else if (th == 32) hipLaunchKernelGGL((k_blur<R2, BASE, 32>), grid, dim3(BL_NT), 0, st, jobs, o, layer, h);
    else hipLaunchKernelGGL((k_blur<R2, BASE, 16>), grid, dim3(BL_NT), 0, st, jobs, o, layer, h);
}

static int launch_layer(const SiftJob* jobs, int n_img, int o, int layer, int R, int C, hipStream_t st) {
    const TapSet& T = taps();
    const HalfTaps& h = T.h[layer];
    switch (T.r2[layer]) {
        case 5: launch_blur<5, false>(jobs, n_img, o, layer, R, C, h, st); break;
        case 6: launch_blur<6, false>(jobs, n_img, o, layer, R, C, h, st); break;
        case 8: launch_blur<8, false>(jobs, n_img, o, layer, R, C, h, st); break;
        case 10: launch_blur<10, false>(jobs, n_img, o, layer, R, C, h, st); break;
        case 13: launch_blur<13, false>(jobs, n_img, o, layer, R, C, h, st); break;
        default:
            set_error("unexpected blur half-width %d", T.r2[layer]);
            return FSX_EINVAL;
    }
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// A run of same-shape images inside the job table (pyramid and extrema grids depend on shape).
struct Group {
    const SiftJob* jobs;
    int n;
    SiftPlan plan;
};

// Enqueue the pipeline for the whole job table: pyramid and extrema per shape group, the
// count-driven stages over all images at once; with `features`, also the grid spread and the
// descriptors of the selection.  ev (optional): 8 events bracketing the stages.
static int run_pipeline(const SiftJob* all, int n_all, const Group* groups, int n_groups, int kp_cap, int nfeat_max,
                        bool features, hipStream_t st, hipEvent_t* ev) {
    const TapSet& T = taps();
    if (T.r2[0] != 5) {
        set_error("unexpected base blur half-width %d", T.r2[0]);
        return FSX_EINVAL;
    }
    if (ev) FSX_HIP(hipEventRecord(ev[0], st));
    hipLaunchKernelGGL(k_reset, dim3(n_all), dim3(64), 0, st, all);
    for (int gi = 0; gi < n_groups; gi++) {
        const Group& g = groups[gi];
        const SiftPlan& p = g.plan;
        launch_blur<5, true>(g.jobs, g.n, 0, 0, p.R[0], p.C[0], T.h[0], st);
        FSX_LAUNCH_CHECK();
        for (int o = 0; o < p.n_oct; o++)
            for (int layer = 1; layer <= N_LAYERS + 2; layer++) {
                int rc = launch_layer(g.jobs, g.n, o, layer, p.R[o], p.C[o], st);
                if (rc) return rc;
            }
    }
    if (ev) FSX_HIP(hipEventRecord(ev[1], st));
    for (int gi = 0; gi < n_groups; gi++) {
        const Group& g = groups[gi];
        const SiftPlan& p = g.plan;
        for (int o = 0; o < p.n_oct; o++) {
            const int strips = (p.C[o] + EX_W - 1) / EX_W;
            dim3 grid((strips + 3) / 4, (p.R[o] + EX_H - 1) / EX_H, g.n);
            hipLaunchKernelGGL(k_extrema, grid, dim3(256), 0, st, g.jobs, o);
            FSX_LAUNCH_CHECK();
        }
    }
    if (ev) FSX_HIP(hipEventRecord(ev[2], st));
    hipLaunchKernelGGL(k_refine, dim3(tune_get(TUNE_REFINE_BLOCKS), 1, n_all), dim3(256), 0, st, all);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[3], st));
    hipLaunchKernelGGL(k_orient, dim3(tune_get(TUNE_ORIENT_BLOCKS), 1, n_all), dim3(64), 0, st, all);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[4], st));
    hipLaunchKernelGGL(k_sort_local, dim3(kp_cap / 1024, 1, n_all), dim3(256), 0, st, all);
    FSX_LAUNCH_CHECK();
    int in0 = 1;
    for (int w = 1024; w < kp_cap; w <<= 1) {
        hipLaunchKernelGGL(k_merge, dim3(std::min(kp_cap / 256, 1024), 1, n_all), dim3(256), 0, st, all, w, in0);
        FSX_LAUNCH_CHECK();
        in0 = !in0;
    }
    hipLaunchKernelGGL(k_emit, dim3(64, 1, n_all), dim3(256), 0, st, all, in0);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[5], st));
    if (features) {
        hipLaunchKernelGGL(k_spread, dim3(1, 1, n_all), dim3(SP_THREADS), 0, st, all);
        FSX_LAUNCH_CHECK();
        if (ev) FSX_HIP(hipEventRecord(ev[6], st));
        if (nfeat_max > 0) {
            hipLaunchKernelGGL(k_describe_sel, dim3(nfeat_max, 1, n_all), dim3(256), 0, st, all);
            FSX_LAUNCH_CHECK();
        }
        if (ev) FSX_HIP(hipEventRecord(ev[7], st));
    }
    return FSX_OK;
}

struct EventSet {
    hipEvent_t ev[8];
    int n = 0;
    ~EventSet() {
        for (int i = 0; i < n; i++) (void)hipEventDestroy(ev[i]);
    }
    int create(int k) {
        for (; n < k; n++) FSX_HIP(hipEventCreate(&ev[n]));
        return FSX_OK;
    }
    int read(float* ms, int k) {
        FSX_HIP(hipEventSynchronize(ev[k]));
        for (int i = 0; i < k; i++) FSX_HIP(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        return FSX_OK;
    }
};

}  // namespace fsx

using namespace fsx;

extern "C" {

int fsx_sift_plan(int H, int W, int* n_oct, int* rows, int* cols) {
    FSX_REQUIRE(H > 0 && W > 0, "bad image size %dx%d", H, W);
    SiftPlan p;
    sift_plan(H, W, &p);
    *n_oct = p.n_oct;
    for (int o = 0; o < p.n_oct; o++) {
        if (rows) rows[o] = p.R[o];
        if (cols) cols[o] = p.C[o];
    }
    return FSX_OK;
}

int fsx_sift_taps(int layer, float* out, int* len) {
    FSX_REQUIRE(layer >= 0 && layer < 6, "layer %d out of range", layer);
    const TapSet& T = taps();
    *len = T.len[layer];
    for (int i = 0; i < T.len[layer]; i++) out[i] = T.full[layer][i];
    return FSX_OK;
}

int fsx_sift_workspace_bytes(int H, int W, int cand_cap, int kp_cap, size_t* bytes) {
    FSX_REQUIRE(H >= 7 && W >= 7, "image %dx%d too small", H, W);
    FSX_REQUIRE(cand_cap > 0 && kp_cap >= 1024 && (kp_cap & (kp_cap - 1)) == 0,
                "kp_cap must be a power of two >= 1024 (got %d)", kp_cap);
    SiftPlan p;
    sift_plan(H, W, &p);
    *bytes = ws_layout(p, cand_cap, kp_cap).total;
    return FSX_OK;
}

int fsx_sift_detect(const uint8_t* img, int H, int W, int ld, double contrast, double edge, int max_kp,
                    int cand_cap, int kp_cap, void* ws, size_t ws_bytes, float* kp_out, int* n_out,
                    float* stage_ms, void* stream) {
    FSX_REQUIRE(img && ws && kp_out && n_out, "null pointer argument");
    FSX_REQUIRE(H >= 7 && W >= 7 && ld >= W, "bad image geometry %dx%d ld %d", H, W, ld);
    FSX_REQUIRE(max_kp >= 1, "max_kp must be >= 1");
    size_t need = 0;
    int rc = fsx_sift_workspace_bytes(H, W, cand_cap, kp_cap, &need);
    if (rc) return rc;
    FSX_REQUIRE(ws_bytes >= need, "workspace too small: %zu < %zu bytes", ws_bytes, need);
    hipStream_t st = (hipStream_t)stream;
    SiftPlan p;
    sift_plan(H, W, &p);
    SiftJob J;
    fill_job(&J, p, img, ld, contrast, edge, max_kp, cand_cap, kp_cap, ws, kp_out, n_out);
    rc = upload_async(ws, &J, sizeof(J), st);
    if (rc) return rc;
    EventSet E;
    if (stage_ms && E.create(6)) return FSX_EHIP;
    Group g{(const SiftJob*)ws, 1, p};
    rc = run_pipeline((const SiftJob*)ws, 1, &g, 1, kp_cap, 0, false, st, stage_ms ? E.ev : nullptr);
    if (rc) return rc;
    if (stage_ms && E.read(stage_ms, 5)) return FSX_EHIP;
    return FSX_OK;
}

int fsx_sift_counters(const void* ws, int* out4, void* stream) {
    FSX_REQUIRE(ws && out4, "null pointer argument");
    hipStream_t st = (hipStream_t)stream;
    SiftJob J;
    FSX_HIP(hipMemcpyAsync(&J, ws, sizeof(J), hipMemcpyDeviceToHost, st));
    FSX_HIP(hipStreamSynchronize(st));
    FSX_HIP(hipMemcpyAsync(out4, J.counters, 4 * sizeof(int), hipMemcpyDeviceToHost, st));
    FSX_HIP(hipStreamSynchronize(st));
    return FSX_OK;
}

int fsx_sift_describe(const void* ws, const float* kp, int kp_stride, const long long* sel, int m, float* desc,
                      uint8_t* u8_out, void* stream) {
    FSX_REQUIRE(ws && kp && desc, "null pointer argument");
    FSX_REQUIRE(m >= 0 && kp_stride >= 1, "bad sizes m=%d stride=%d", m, kp_stride);
    if (m == 0) return FSX_OK;
    FSX_REQUIRE(sel != nullptr, "null selection");
    hipStream_t st = (hipStream_t)stream;
    DescArgs A{kp, kp_stride, sel, m, desc, u8_out};
    hipLaunchKernelGGL(k_describe, dim3(m), dim3(256), 0, st, (const SiftJob*)ws, A);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_sift_table_bytes(int n_img, size_t* bytes) {
    FSX_REQUIRE(n_img >= 1, "n_img must be >= 1");
    *bytes = (size_t)n_img * sizeof(SiftJob);
    return FSX_OK;
}

}  // extern "C"

namespace fsx {

// host-side job records and shape groups of a feature batch (images grouped by shape; the
// device table holds each group contiguously)
static int build_feature_jobs(int n_img, const uint8_t* const* imgs, const int* H, const int* W, const int* ld,
                              double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                              void* const* ws, const size_t* ws_bytes, float* const* kp_out, int* n_out_dev,
                              float* const* desc_out, uint8_t* const* u8_out, void* table_dev, size_t table_bytes,
                              std::vector<SiftJob>& jobs, std::vector<int>& order, std::vector<Group>& groups) {
    FSX_REQUIRE(n_img >= 1 && imgs && H && W && ld && ws && ws_bytes && kp_out && n_out_dev && table_dev,
                "null pointer argument");
    FSX_REQUIRE(max_kp >= 1 && nfeat >= 0 && max_kp <= kp_cap, "need 1 <= max_kp <= kp_cap (max_kp %d kp_cap %d)",
                max_kp, kp_cap);
    FSX_REQUIRE(table_bytes >= (size_t)n_img * sizeof(SiftJob), "job table too small");
    FSX_REQUIRE(max_kp < 65536, "max_kp must be < 65536 for the device spread (got %d)", max_kp);
    order.resize(n_img);
    for (int i = 0; i < n_img; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return H[a] != H[b] ? H[a] < H[b] : W[a] < W[b];
    });
    jobs.assign(n_img, SiftJob{});
    groups.clear();
    SiftJob* dev = (SiftJob*)table_dev;
    for (int k = 0; k < n_img; k++) {
        const int i = order[k];
        FSX_REQUIRE(imgs[i] && ws[i] && kp_out[i], "null pointer for image %d", i);
        FSX_REQUIRE(H[i] >= 7 && W[i] >= 7 && ld[i] >= W[i], "bad geometry for image %d", i);
        size_t need = 0;
        int rc = fsx_sift_workspace_bytes(H[i], W[i], cand_cap, kp_cap, &need);
        if (rc) return rc;
        FSX_REQUIRE(ws_bytes[i] >= need, "workspace %d too small: %zu < %zu bytes", i, ws_bytes[i], need);
        SiftPlan p;
        sift_plan(H[i], W[i], &p);
        SiftJob& J = jobs[k];
        fill_job(&J, p, imgs[i], ld[i], contrast, edge, max_kp, cand_cap, kp_cap, ws[i], kp_out[i], n_out_dev + i);
        J.nfeat = nfeat;
        J.desc_out = desc_out ? desc_out[i] : nullptr;
        J.u8_out = u8_out ? u8_out[i] : nullptr;
        if (groups.empty() || groups.back().plan.H != H[i] || groups.back().plan.W != W[i])
            groups.push_back(Group{dev + k, 0, p});
        groups.back().n++;
    }
    return FSX_OK;
}

}  // namespace fsx

extern "C" {

int fsx_sift_features_prepare(int n_img, const uint8_t* const* imgs, const int* H, const int* W, const int* ld,
                              double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                              void* const* ws, const size_t* ws_bytes, float* const* kp_out, int* n_out_dev,
                              float* const* desc_out, uint8_t* const* u8_out, void* table_dev, size_t table_bytes,
                              void* stream) {
    std::vector<SiftJob> jobs;
    std::vector<int> order;
    std::vector<Group> groups;
    int rc = build_feature_jobs(n_img, imgs, H, W, ld, contrast, edge, max_kp, nfeat, cand_cap, kp_cap, ws, ws_bytes,
                                kp_out, n_out_dev, desc_out, u8_out, table_dev, table_bytes, jobs, order, groups);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    for (int k = 0; k < n_img; k++) {
        rc = upload_async(ws[order[k]], &jobs[k], sizeof(SiftJob), st);
        if (rc) return rc;
    }
    return upload_async(table_dev, jobs.data(), (size_t)n_img * sizeof(SiftJob), st);
}

int fsx_sift_features_launch(int n_img, const uint8_t* const* imgs, const int* H, const int* W, const int* ld,
                             double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                             void* const* ws, const size_t* ws_bytes, float* const* kp_out, int* n_out_dev,
                             float* const* desc_out, uint8_t* const* u8_out, void* table_dev, size_t table_bytes,
                             void* const* stage_events, void* stream) {
    std::vector<SiftJob> jobs;
    std::vector<int> order;
    std::vector<Group> groups;
    int rc = build_feature_jobs(n_img, imgs, H, W, ld, contrast, edge, max_kp, nfeat, cand_cap, kp_cap, ws, ws_bytes,
                                kp_out, n_out_dev, desc_out, u8_out, table_dev, table_bytes, jobs, order, groups);
    if (rc) return rc;
    hipEvent_t ev[8];
    if (stage_events)
        for (int i = 0; i < 8; i++) ev[i] = (hipEvent_t)stage_events[i];
    return run_pipeline((const SiftJob*)table_dev, n_img, groups.data(), (int)groups.size(), kp_cap, nfeat, true,
                        (hipStream_t)stream, stage_events ? ev : nullptr);
}

int fsx_sift_features_batch(int n_img, const uint8_t* const* imgs, const int* H, const int* W, const int* ld,
                            double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                            void* const* ws, const size_t* ws_bytes, float* const* kp_out, int* n_out_dev,
                            float* const* desc_out, uint8_t* const* u8_out, void* table_dev, size_t table_bytes,
                            void* const* stage_events, void* stream) {
    int rc = fsx_sift_features_prepare(n_img, imgs, H, W, ld, contrast, edge, max_kp, nfeat, cand_cap, kp_cap, ws,
                                       ws_bytes, kp_out, n_out_dev, desc_out, u8_out, table_dev, table_bytes, stream);
    if (rc) return rc;
    return fsx_sift_features_launch(n_img, imgs, H, W, ld, contrast, edge, max_kp, nfeat, cand_cap, kp_cap, ws,
                                    ws_bytes, kp_out, n_out_dev, desc_out, u8_out, table_dev, table_bytes,
                                    stage_events, stream);
}

// Largest atan2f error in float32 ulps over n seeded pairs (k_atan2_probe); *max_ulp_dev must
// be zero on entry.
int fsx_atan2_probe(long long n, unsigned long long seed, unsigned* max_ulp_dev, void* stream) {
    FSX_REQUIRE(n >= 1 && max_ulp_dev, "bad atan2 probe arguments");
    hipLaunchKernelGGL(k_atan2_probe, dim3(1024), dim3(256), 0, (hipStream_t)stream, n, seed, max_ulp_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Largest error of the descriptor's gradient angle and of the atan2f form, in 1e-9 degree,
// over n seeded pairs (k_desc_angle_probe); out[0..1] must hold zeros on entry.
int fsx_desc_angle_probe(long long n, unsigned long long seed, unsigned* out, void* stream) {
    FSX_REQUIRE(n >= 1 && out, "bad descriptor angle probe arguments");
    hipLaunchKernelGGL(k_desc_angle_probe, dim3(1024), dim3(256), 0, (hipStream_t)stream, n, seed, out);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_sift_selection_offset(int H, int W, int cand_cap, int kp_cap, size_t* off) {
    size_t total;
    int rc = fsx_sift_workspace_bytes(H, W, cand_cap, kp_cap, &total);
    if (rc) return rc;
    SiftPlan p;
    sift_plan(H, W, &p);
    *off = ws_layout(p, cand_cap, kp_cap).sel;
    return FSX_OK;
}

}  // extern "C"
