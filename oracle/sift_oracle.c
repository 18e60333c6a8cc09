/* CPU oracle of the SIFT front end: a plain-C restatement of the reference's CPU path
 * (app/planesite/gsift.py, numpy/scipy semantics), used ONLY as test infrastructure: as the
 * checker in tests/, smoke() and the CPU-baseline leg of bench.py.  Never as a product path.
 *
 * Arithmetic follows the reference's numpy dtypes operation by operation:
 *   - float32 everywhere numpy computes in float32 (weak Python scalars become float32),
 *     compiled with -ffp-contract=off so no fused multiply-add changes a rounding;
 *   - float64 where numpy promotes (scipy.ndimage.correlate1d accumulates in float64 with
 *     the symmetric-pair order; bincount accumulates in float64; the descriptor's sample
 *     geometry is float64 because the host-side scale factors are float64);
 *   - numpy's own reduction order (pairwise sum with 8 accumulators for 8 <= n <= 128,
 *     left to right below 8).
 * Transcendentals (exp, cos, sin, atan2, pow) are evaluated in float64 and rounded once
 * (correctly rounded float32 results); numpy's float32 SIMD versions are not correctly
 * rounded and depend on the host CPU, so those few ulps are a tolerance, not a bit match.
 *
 * Parity pins (tests/test_oracle_*.py): golden vectors written from the reference's CPU path
 * (tests/golden/), scipy.ndimage.correlate1d / maximum_filter, numpy.linalg.solve on analytic
 * quadratic volumes, synthetic oriented gradient fields and brute force.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define N_LAYERS 3
#define SIGMA 1.6
#define INIT_SIGMA 0.5
#define BORDER 5
#define MAX_INTERP 5
#define ORI_BINS 36
#define ORI_RADIUS 4.5
#define ORI_SIG 1.5
#define ORI_PEAK 0.8
#define D_WIDTH 4
#define D_BINS 8
#define D_SCL 3.0
#define D_MAG_THR 0.2
#define MAXO 16

/* numpy's pairwise summation (numpy/_core/src/umath/loops_utils.h.src) for n <= 128, float64 */
static double np_sum_d(const double *a, int n) {
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

/* the same for float32 (the 128-bin descriptor norms) */
static float np_sum_f(const float *a, int n) {
    if (n < 8) {
        float r = 0.0f;
        for (int i = 0; i < n; i++) r += a[i];
        return r;
    }
    float r[8];
    for (int j = 0; j < 8; j++) r[j] = a[j];
    int i;
    for (i = 8; i < n - (n % 8); i += 8)
        for (int j = 0; j < 8; j++) r[j] += a[i + j];
    float res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; i++) res += a[i];
    return res;
}

static inline float exp_f(float x) { return (float)exp((double)x); }
static inline float cos_f(float x) { return (float)cos((double)x); }
static inline float sin_f(float x) { return (float)sin((double)x); }
static inline float atan2_f(float y, float x) { return (float)atan2((double)y, (double)x); }

/* gsift.py:52-56 _gauss_kernel: round(8 sigma + 1) | 1 taps, float64 exp, normalised, float32 */
int fso_gauss_kernel(double sigma, float *out) {
    int k = ((int)rint(sigma * 8.0 + 1.0)) | 1;
    double g[64];
    if (k > 64) return -1;
    for (int i = 0; i < k; i++) {
        double x = (double)i - (k - 1) / 2.0;
        g[i] = exp(-x * x / (2.0 * sigma * sigma));
    }
    double s = np_sum_d(g, k);
    for (int i = 0; i < k; i++) out[i] = (float)(g[i] / s);
    return k;
}

/* gsift.py:59-72 _Kernels: sigma[0] = base blur, sigma[1..5] = layer increments */
void fso_sigmas(double *sig6) {
    double k = pow(2.0, 1.0 / N_LAYERS);
    sig6[0] = sqrt(fmax(SIGMA * SIGMA - (2 * INIT_SIGMA) * (2 * INIT_SIGMA), 0.01));
    for (int i = 1; i < N_LAYERS + 3; i++) {
        double prev = pow(k, i - 1) * SIGMA;
        double total = prev * k;
        sig6[i] = sqrt(total * total - prev * prev);
    }
}

/* gsift.py:167-172 octave count and the min(shape) < 2*BORDER+4 stop; [::2] halves round up */
int fso_plan(int H, int W, int *R, int *C) {
    int mn = H < W ? H : W;
    int n_oct = (int)rint(log2((double)mn) - 2.0) + 1;
    if (n_oct < 1) n_oct = 1;
    int r = 2 * H, c = 2 * W, n = 0;
    for (int o = 0; o < n_oct && o < MAXO; o++) {
        if ((r < c ? r : c) < 2 * BORDER + 4) break;
        R[n] = r;
        C[n] = c;
        n++;
        r = (r + 1) / 2;
        c = (c + 1) / 2;
    }
    return n;
}

static inline int reflect101(int p, int n) {
    if (n == 1) return 0;
    while (p < 0 || p >= n) p = p < 0 ? -p : 2 * n - 2 - p;
    return p;
}

/* scipy.ndimage.correlate1d(mode="mirror") for an odd symmetric float32 kernel: float64
 * accumulation, centre tap first, then the symmetric pairs from the outermost inwards
 * (scipy/ndimage/src/ni_filters.c NI_Correlate1D, symmetric branch); output rounded to f32.
 * gsift.py:85-87 _blur = axis 0 then axis 1. */
void fso_blur(const float *in, float *out, int R, int C, const float *k, int klen) {
    int r2 = klen / 2;
    float *t = (float *)malloc(sizeof(float) * (size_t)R * C);
    double kd[64];
    for (int j = 0; j < klen; j++) kd[j] = (double)k[j];
    /* axis 0 */
    for (int c = 0; c < C; c++) {
        for (int r = 0; r < R; r++) {
            double s = (double)in[(size_t)r * C + c] * kd[r2];
            for (int j = r2; j >= 1; j--) {
                double a = (double)in[(size_t)reflect101(r - j, R) * C + c];
                double b = (double)in[(size_t)reflect101(r + j, R) * C + c];
                s += (a + b) * kd[r2 - j];
            }
            t[(size_t)r * C + c] = (float)s;
        }
    }
    /* axis 1 */
    for (int r = 0; r < R; r++) {
        const float *row = t + (size_t)r * C;
        for (int c = 0; c < C; c++) {
            double s = (double)row[c] * kd[r2];
            for (int j = r2; j >= 1; j--) {
                double a = (double)row[reflect101(c - j, C)];
                double b = (double)row[reflect101(c + j, C)];
                s += (a + b) * kd[r2 - j];
            }
            out[(size_t)r * C + c] = (float)s;
        }
    }
    free(t);
}

/* CPU model of the device's fused blur evaluation (fieldstitch_amd/csrc/sift.hip,
 * blur_point; not part of the reference).  For n windows of 2*r2+1 float32 values and the
 * half taps kh[0..r2] (centre first), the reference order (as in fso_blur above) is compared
 * with the fused form acc = fma(w[-j] + w[+j], k[j], acc).  use_check 1: the fused value is
 * accepted only when acc*(1-margin) and acc*(1+margin) round to the same float32; use_check 2
 * (the device's test): accepted unless acc's 29 mantissa bits below float32 precision lie within
 * `margin` float64 ulps of the midpoint pattern 0x10000000, or 0 < acc < 2^-120.  Otherwise the
 * reference order is evaluated (counted in *n_fallback).  Returns the number of windows whose
 * float32 result differs from the reference's. */
int64_t fso_fused_blur_check(const float *win, int64_t n, int r2, const float *kh, double margin, int use_check,
                             int64_t *n_fallback) {
    double k[64];
    for (int j = 0; j <= r2; j++) k[j] = (double)kh[j];
    int64_t bad = 0, fb = 0;
    for (int64_t i = 0; i < n; i++) {
        const float *w = win + i * (2 * r2 + 1);
        double ref = (double)w[r2] * k[0];
        for (int j = r2; j >= 1; j--) ref += ((double)w[r2 - j] + (double)w[r2 + j]) * k[j];
        double acc = (double)w[r2] * k[0];
        for (int j = r2; j >= 1; j--) acc = fma((double)w[r2 - j] + (double)w[r2 + j], k[j], acc);
        float out;
        if (use_check == 2) {
            uint64_t bits;
            memcpy(&bits, &acc, 8);
            const uint32_t low29 = (uint32_t)bits & 0x1fffffffu, m = (uint32_t)margin;
            const int near_mid = (uint32_t)(low29 - (0x10000000u - m)) <= 2u * m;
            /* the device tests the float32 result: 0 < r < 2^-120 */
            const float r = (float)acc;
            uint32_t rb;
            memcpy(&rb, &r, 4);
            const int tiny = rb - 1u < 0x037FFFFFu;
            if (!(near_mid || tiny)) {
                out = (float)acc;
            } else {
                out = (float)ref;
                fb++;
            }
        } else if (use_check) {
            const float lo = (float)(acc * (1.0 - margin)), hi = (float)(acc * (1.0 + margin));
            if (lo == hi) {
                out = lo;
            } else {
                out = (float)ref;
                fb++;
            }
        } else {
            out = (float)acc;
        }
        bad += out != (float)ref;
    }
    if (n_fallback) *n_fallback = fb;
    return bad;
}

/* gsift.py:90-102 _up2 on img = u8 * float32(1/255) (gsift.py:160): rows then columns */
void fso_up2_u8(const uint8_t *img, int H, int W, int ld, float *out) {
    const float s255 = (float)(1.0 / 255.0);
    float *a = (float *)malloc(sizeof(float) * (size_t)H * W);
    float *t = (float *)malloc(sizeof(float) * (size_t)2 * H * W);
    for (int r = 0; r < H; r++)
        for (int c = 0; c < W; c++) a[(size_t)r * W + c] = (float)img[(size_t)r * ld + c] * s255;
    for (int y = 0; y < 2 * H; y++) {
        int iy = y >> 1;
        int ny = (y & 1) ? (iy + 1 < H ? iy + 1 : H - 1) : (iy - 1 > 0 ? iy - 1 : 0);
        for (int c = 0; c < W; c++) {
            float p = 0.75f * a[(size_t)iy * W + c];
            float q = 0.25f * a[(size_t)ny * W + c];
            t[(size_t)y * W + c] = p + q;
        }
    }
    for (int y = 0; y < 2 * H; y++)
        for (int x = 0; x < 2 * W; x++) {
            int ix = x >> 1;
            int nx = (x & 1) ? (ix + 1 < W ? ix + 1 : W - 1) : (ix - 1 > 0 ? ix - 1 : 0);
            float p = 0.75f * t[(size_t)y * W + ix];
            float q = 0.25f * t[(size_t)y * W + nx];
            out[(size_t)y * 2 * W + x] = p + q;
        }
    free(a);
    free(t);
}

typedef struct {
    int n_oct;
    int R[MAXO], C[MAXO];
    float *G[MAXO][N_LAYERS + 3];   /* gaussian layers */
    float *D[MAXO];                 /* (5, R, C) DoG */
    float *MAG[MAXO], *ANG[MAXO];   /* (3, R, C) gradient of gauss layers 1..3 */
    int64_t dbase[MAXO], gbase[MAXO];
} fso_sift;

void fso_sift_free(fso_sift *s) {
    if (!s) return;
    for (int o = 0; o < s->n_oct; o++) {
        for (int i = 0; i < N_LAYERS + 3; i++) free(s->G[o][i]);
        free(s->D[o]);
        free(s->MAG[o]);
        free(s->ANG[o]);
    }
    free(s);
}

/* gsift.py:158-179 _pyramid and gsift.py:227-267 _Atlas (CPU branch) */
fso_sift *fso_sift_new(const uint8_t *img, int H, int W, int ld) {
    fso_sift *s = (fso_sift *)calloc(1, sizeof(fso_sift));
    double sig[6];
    float taps[6][64];
    int klen[6];
    fso_sigmas(sig);
    for (int i = 0; i < 6; i++) klen[i] = fso_gauss_kernel(sig[i], taps[i]);
    s->n_oct = fso_plan(H, W, s->R, s->C);
    float *up = (float *)malloc(sizeof(float) * (size_t)4 * H * W);
    fso_up2_u8(img, H, W, ld, up);
    float *cur = (float *)malloc(sizeof(float) * (size_t)4 * H * W);
    fso_blur(up, cur, 2 * H, 2 * W, taps[0], klen[0]);
    free(up);
    int64_t db = 0, gb = 0;
    for (int o = 0; o < s->n_oct; o++) {
        int R = s->R[o], C = s->C[o];
        size_t P = (size_t)R * C;
        s->G[o][0] = cur;
        for (int i = 1; i < N_LAYERS + 3; i++) {
            s->G[o][i] = (float *)malloc(sizeof(float) * P);
            fso_blur(s->G[o][i - 1], s->G[o][i], R, C, taps[i], klen[i]);
        }
        s->D[o] = (float *)malloc(sizeof(float) * 5 * P);
        for (int i = 0; i < N_LAYERS + 2; i++)
            for (size_t p = 0; p < P; p++) s->D[o][i * P + p] = s->G[o][i + 1][p] - s->G[o][i][p];
        s->MAG[o] = (float *)malloc(sizeof(float) * 3 * P);
        s->ANG[o] = (float *)malloc(sizeof(float) * 3 * P);
        const float deg = 180.0f / 3.14159265358979323846f;   /* numpy degrees(): x * (180f / pi_f) */
        for (int li = 1; li <= N_LAYERS; li++) {
            const float *G = s->G[o][li];
            float *M = s->MAG[o] + (li - 1) * P, *A = s->ANG[o] + (li - 1) * P;
            for (int r = 0; r < R; r++)
                for (int c = 0; c < C; c++) {
                    size_t k = (size_t)r * C + c;
                    float dx = 0.f, dy = 0.f;
                    if (r > 0 && r < R - 1 && c > 0 && c < C - 1) {
                        dx = G[k + 1] - G[k - 1];
                        dy = G[k - C] - G[k + C];
                    }
                    float xx = dx * dx, yy = dy * dy;
                    M[k] = sqrtf(xx + yy);
                    float a = atan2_f(dy, dx) * deg;
                    A[k] = a < 0.f ? a + 360.0f : a;
                }
        }
        s->dbase[o] = db;
        s->gbase[o] = gb;
        db += 5 * (int64_t)P;
        gb += 3 * (int64_t)P;
        if (o + 1 < s->n_oct) {
            int R2 = s->R[o + 1], C2 = s->C[o + 1];
            cur = (float *)malloc(sizeof(float) * (size_t)R2 * C2);
            const float *G3 = s->G[o][N_LAYERS];
            for (int r = 0; r < R2; r++)
                for (int c = 0; c < C2; c++) cur[(size_t)r * C2 + c] = G3[(size_t)(2 * r) * C + 2 * c];
        }
    }
    return s;
}

int fso_sift_n_oct(const fso_sift *s) { return s->n_oct; }
void fso_sift_octave(const fso_sift *s, int o, int *R, int *C) { *R = s->R[o]; *C = s->C[o]; }
const float *fso_sift_gauss(const fso_sift *s, int o, int i) { return s->G[o][i]; }
const float *fso_sift_dog(const fso_sift *s, int o) { return s->D[o]; }
const float *fso_sift_mag(const fso_sift *s, int o) { return s->MAG[o]; }
const float *fso_sift_ang(const fso_sift *s, int o) { return s->ANG[o]; }

static inline float dog_at(const fso_sift *s, int o, int l, int r, int c) {
    return s->D[o][((size_t)l * s->R[o] + r) * s->C[o] + c];
}

/* gsift.py:293-319 _extrema (CPU branch): non-strict 3x3x3 max/min with the prefilter,
 * layers 1..3 and 5-pixel border; returns ascending flat DoG-atlas indices */
int64_t fso_sift_extrema(const fso_sift *s, double contrast, int64_t *out, int64_t cap) {
    const float pthr = (float)(0.8 * 0.5 * contrast / N_LAYERS);
    int64_t n = 0;
    for (int o = 0; o < s->n_oct; o++) {
        int R = s->R[o], C = s->C[o];
        for (int l = 1; l <= N_LAYERS; l++)
            for (int r = BORDER; r < R - BORDER; r++)
                for (int c = BORDER; c < C - BORDER; c++) {
                    float v = dog_at(s, o, l, r, c);
                    int mx = v > pthr, mn = v < -pthr;
                    if (!mx && !mn) continue;
                    for (int dl = -1; dl <= 1 && (mx || mn); dl++)
                        for (int dr = -1; dr <= 1; dr++)
                            for (int dc = -1; dc <= 1; dc++) {
                                float u = dog_at(s, o, l + dl, r + dr, c + dc);
                                if (u > v) mx = 0;
                                if (u < v) mn = 0;
                            }
                    if (mx || mn) {
                        if (n < cap) out[n] = s->dbase[o] + ((int64_t)l * R + r) * C + c;
                        n++;
                    }
                }
    }
    return n;
}

typedef struct { float v, dx, dy, ds, dxx, dyy, dss, dxy, dxs, dys; } derivs_t;

/* gsift.py:274-290 _derivs_flat */
static derivs_t derivs(const fso_sift *s, int o, int l, int r, int c) {
    derivs_t d;
    const float *D = s->D[o];
    int64_t C = s->C[o], RC = (int64_t)s->R[o] * C;
    int64_t k = l * RC + r * C + c;
    d.v = D[k];
    float xp1 = D[k + 1], xm1 = D[k - 1], yp1 = D[k + C], ym1 = D[k - C], sp1 = D[k + RC], sm1 = D[k - RC];
    d.dx = (xp1 - xm1) * 0.5f;
    d.dy = (yp1 - ym1) * 0.5f;
    d.ds = (sp1 - sm1) * 0.5f;
    d.dxx = (xp1 + xm1) - 2.0f * d.v;
    d.dyy = (yp1 + ym1) - 2.0f * d.v;
    d.dss = (sp1 + sm1) - 2.0f * d.v;
    d.dxy = (((D[k + C + 1] - D[k + C - 1]) - D[k - C + 1]) + D[k - C - 1]) * 0.25f;
    d.dxs = (((D[k + RC + 1] - D[k + RC - 1]) - D[k - RC + 1]) + D[k - RC - 1]) * 0.25f;
    d.dys = (((D[k + RC + C] - D[k + RC - C]) - D[k - RC + C]) + D[k - RC - C]) * 0.25f;
    return d;
}

/* gsift.py:182-198 _inv3_solve for the symmetric Hessian, then x = -(inv g); returns 0 if |det| <= 1e-20 */
static int solve3(const derivs_t *d, float *x) {
    float a = d->dxx, bb = d->dxy, c = d->dxs, dd = d->dxy, e = d->dyy, f = d->dys, g = d->dxs, h = d->dys, i = d->dss;
    float A = e * i - f * h;
    float B = -(dd * i - f * g);
    float Cc = dd * h - e * g;
    float det = (a * A + bb * B) + c * Cc;
    int ok = fabsf(det) > 1e-20f;
    if (!ok) det = 1.0f;
    float inv[3][3] = {
        {A / det, -(bb * i - c * h) / det, (bb * f - c * e) / det},
        {B / det, (a * i - c * g) / det, -(a * f - c * dd) / det},
        {Cc / det, -(a * h - bb * g) / det, (a * e - bb * dd) / det}};
    float b3[3] = {d->dx, d->dy, d->ds};
    for (int r = 0; r < 3; r++) x[r] = -((inv[r][0] * b3[0] + inv[r][1] * b3[1]) + inv[r][2] * b3[2]);
    return ok;
}

/* gsift.py:322-369 _refine (CPU branch, vectorised semantics per candidate).  Output keeps
 * candidate order.  Returns the number kept. */
int64_t fso_sift_refine(const fso_sift *s, const int64_t *F, int64_t n, double contrast, double edge,
                        int32_t *oo, int32_t *ol, int32_t *orr, int32_t *oc, float *oX, float *oresp,
                        int64_t *osrc) {
    const float contrast_f = (float)contrast, edge_f = (float)edge;
    const float e1 = (float)((edge + 1.0) * (edge + 1.0));
    int64_t m = 0;
    for (int64_t i = 0; i < n; i++) {
        int o = 0;
        while (o + 1 < s->n_oct && F[i] >= s->dbase[o + 1]) o++;
        int R = s->R[o], C = s->C[o];
        int64_t RC = (int64_t)R * C, loc = F[i] - s->dbase[o];
        int64_t l = loc / RC, rem = loc - l * RC, r = rem / C, c = rem - r * C;
        int alive = 1, done = 0;
        float X[3] = {0, 0, 0};
        for (int it = 0; it < MAX_INTERP; it++) {
            derivs_t d = derivs(s, o, (int)l, (int)r, (int)c);
            float x[3];
            int okd = solve3(&d, x);
            alive &= okd;
            int conv = fabsf(x[0]) < 0.5f && fabsf(x[1]) < 0.5f && fabsf(x[2]) < 0.5f;
            if (alive && !done && conv) {
                X[0] = x[0]; X[1] = x[1]; X[2] = x[2];
                done = 1;
            }
            int move = alive && !done;
            int big = fabsf(x[0]) > 1e6f || fabsf(x[1]) > 1e6f || fabsf(x[2]) > 1e6f;
            if (move && big) alive = 0;
            move = alive && !done;
            if (move) {
                /* NaN steps (never "big") leave the voxel range: the reference's int cast of
                 * rint(NaN) is far out of range, so the inside test below rejects them */
                if (isnan(x[0]) || isnan(x[1]) || isnan(x[2])) { alive = 0; }
                else {
                    c += (int64_t)rintf(x[0]);
                    r += (int64_t)rintf(x[1]);
                    l += (int64_t)rintf(x[2]);
                }
            }
            int inside = l >= 1 && l <= N_LAYERS && c >= BORDER && c < C - BORDER && r >= BORDER && r < R - BORDER;
            alive &= inside;
            if (l < 1) l = 1;
            if (l > N_LAYERS) l = N_LAYERS;
            if (r < BORDER) r = BORDER;
            if (r > R - BORDER - 1) r = R - BORDER - 1;
            if (c < BORDER) c = BORDER;
            if (c > C - BORDER - 1) c = C - BORDER - 1;
        }
        if (!(alive && done)) continue;
        derivs_t d = derivs(s, o, (int)l, (int)r, (int)c);
        float gX = (d.dx * X[0] + d.dy * X[1]) + d.ds * X[2];
        float contr = d.v + 0.5f * gX;
        float tr = d.dxx + d.dyy;
        float det = d.dxx * d.dyy - d.dxy * d.dxy;
        float ac = fabsf(contr);
        if (ac * 3.0f >= contrast_f && det > 0.0f && (tr * tr) * edge_f < e1 * det) {
            oo[m] = o; ol[m] = (int32_t)l; orr[m] = (int32_t)r; oc[m] = (int32_t)c;
            oX[3 * m] = X[0]; oX[3 * m + 1] = X[1]; oX[3 * m + 2] = X[2];
            oresp[m] = ac;
            if (osrc) osrc[m] = i;
            m++;
        }
    }
    return m;
}

/* gsift.py:372-412 orientation histogram of one keypoint (float64 bincount in window order) */
void fso_sift_ori_hist(const fso_sift *s, int o, int li, int r, int c, float scl, float *hist) {
    int R = s->R[o], C = s->C[o];
    int RM = (int)rint(ORI_RADIUS * SIGMA * pow(2.0, (N_LAYERS + 0.5) / N_LAYERS));   /* ORI_R_MAX = 16 */
    int rad = (int)rintf(4.5f * scl);
    float sw = 1.5f * scl;
    float den = 2.0f * (sw * sw);
    double h[ORI_BINS];
    for (int b = 0; b < ORI_BINS; b++) h[b] = 0.0;
    const float *M = s->MAG[o] + (size_t)(li - 1) * R * C, *A = s->ANG[o] + (size_t)(li - 1) * R * C;
    for (int i = -RM; i <= RM; i++)
        for (int j = -RM; j <= RM; j++) {
            int y = r + i, x = c + j;
            int valid = abs(i) <= rad && abs(j) <= rad && y > 0 && y < R - 1 && x > 0 && x < C - 1;
            if (!valid) continue;
            float d2 = (float)(i * i + j * j);
            float w = exp_f(-d2 / den);
            size_t k = (size_t)y * C + x;
            float m = M[k] * w;
            int b = ((int)rintf(A[k] * (float)(ORI_BINS / 360.0))) % ORI_BINS;
            if (b < 0) b += ORI_BINS;
            h[b] += (double)m;
        }
    for (int b = 0; b < ORI_BINS; b++) hist[b] = (float)h[b];
}

/* gsift.py:415-432 _ori_peaks for one histogram; returns the number of peaks, angles ascending by bin */
int fso_ori_peaks(const float *hist, float *angles) {
    float h[ORI_BINS];
    for (int j = 0; j < ORI_BINS; j++) {
        float a = hist[(j + ORI_BINS - 2) % ORI_BINS] + hist[(j + 2) % ORI_BINS];
        float b = hist[(j + ORI_BINS - 1) % ORI_BINS] + hist[(j + 1) % ORI_BINS];
        h[j] = (a * 0.0625f + b * 0.25f) + hist[j] * 0.375f;
    }
    float hmax = h[0];
    for (int j = 1; j < ORI_BINS; j++) hmax = h[j] > hmax ? h[j] : hmax;
    const float thr = 0.8f * hmax;
    int n = 0;
    for (int j = 0; j < ORI_BINS; j++) {
        float hl = h[(j + ORI_BINS - 1) % ORI_BINS], hc = h[j], hr = h[(j + 1) % ORI_BINS];
        if (!(hc > hl && hc > hr && hc >= thr && hmax > 0.0f)) continue;
        float den = (hl - 2.0f * hc) + hr;
        if (fabsf(den) < 1e-12f) den = -1e-12f;
        float q = (0.5f * (hl - hr)) / den;
        double bin = (double)j + (double)q;
        if (bin < 0) bin += ORI_BINS;
        if (bin >= ORI_BINS) bin -= ORI_BINS;
        double angle = 360.0 - (360.0 / ORI_BINS) * bin;
        if (fabs(angle - 360.0) < 1e-4) angle = 0.0;
        angles[n++] = (float)angle;
    }
    return n;
}

typedef struct { float resp; int64_t seq; } sortkey_t;
static int cmp_key(const void *a, const void *b) {
    const sortkey_t *x = (const sortkey_t *)a, *y = (const sortkey_t *)b;
    /* argsort(-response, kind="stable") */
    float nx = -x->resp, ny = -y->resp;
    if (nx < ny) return -1;
    if (nx > ny) return 1;
    return x->seq < y->seq ? -1 : (x->seq > y->seq ? 1 : 0);
}

/* gsift.py:435-477 detect.  Outputs at most max_kp keypoints (strongest first).  Returns the
 * number written, or -(number needed) if an internal buffer of `cap` entries is too small. */
int64_t fso_sift_detect(const fso_sift *s, double contrast, double edge, int64_t max_kp,
                        float *x, float *y, float *size, float *angle, float *resp,
                        int32_t *octave, int32_t *layer) {
    int64_t capF = 1 << 16;
    int64_t *F = NULL;
    int64_t nF;
    for (;;) {
        F = (int64_t *)realloc(F, sizeof(int64_t) * capF);
        nF = fso_sift_extrema(s, contrast, F, capF);
        if (nF <= capF) break;
        capF = nF;
    }
    int32_t *oo = malloc(4 * (nF + 1)), *ol = malloc(4 * (nF + 1)), *orr = malloc(4 * (nF + 1)), *oc = malloc(4 * (nF + 1));
    float *oX = malloc(12 * (nF + 1)), *ores = malloc(4 * (nF + 1));
    int64_t n = fso_sift_refine(s, F, nF, contrast, edge, oo, ol, orr, oc, oX, ores, NULL);
    /* orientations: every peak becomes a keypoint (kp ascending, bin ascending) */
    int64_t capK = 4 * n + 16, nk = 0;
    float *kx = malloc(4 * capK), *ky = malloc(4 * capK), *ks = malloc(4 * capK), *ka = malloc(4 * capK), *kr = malloc(4 * capK);
    int32_t *ko = malloc(4 * capK), *kl = malloc(4 * capK);
    for (int64_t i = 0; i < n; i++) {
        float e = ((float)ol[i] + oX[3 * i + 2]) / 3.0f;
        float scl = 1.6f * (float)pow(2.0, (double)e);
        float hist[ORI_BINS], ang[ORI_BINS];
        fso_sift_ori_hist(s, oo[i], ol[i], orr[i], oc[i], scl, hist);
        int np_ = fso_ori_peaks(hist, ang);
        float f = (float)ldexp(1.0, oo[i]) * 0.5f;
        for (int p = 0; p < np_; p++) {
            if (nk >= capK) {
                capK *= 2;
                kx = realloc(kx, 4 * capK); ky = realloc(ky, 4 * capK); ks = realloc(ks, 4 * capK);
                ka = realloc(ka, 4 * capK); kr = realloc(kr, 4 * capK); ko = realloc(ko, 4 * capK);
                kl = realloc(kl, 4 * capK);
            }
            kx[nk] = ((float)oc[i] + oX[3 * i]) * f;
            ky[nk] = ((float)orr[i] + oX[3 * i + 1]) * f;
            ks[nk] = (scl * 2.0f) * f;
            ka[nk] = ang[p];
            kr[nk] = ores[i];
            ko[nk] = oo[i];
            kl[nk] = ol[i];
            nk++;
        }
    }
    sortkey_t *keys = malloc(sizeof(sortkey_t) * (nk + 1));
    for (int64_t i = 0; i < nk; i++) { keys[i].resp = kr[i]; keys[i].seq = i; }
    qsort(keys, nk, sizeof(sortkey_t), cmp_key);
    int64_t nout = nk < max_kp ? nk : max_kp;
    for (int64_t i = 0; i < nout; i++) {
        int64_t k = keys[i].seq;
        x[i] = kx[k]; y[i] = ky[k]; size[i] = ks[k]; angle[i] = ka[k]; resp[i] = kr[k];
        octave[i] = ko[k]; layer[i] = kl[k];
    }
    free(keys); free(kx); free(ky); free(ks); free(ka); free(kr); free(ko); free(kl);
    free(F); free(oo); free(ol); free(orr); free(oc); free(oX); free(ores);
    return nout;
}

/* gsift.py:480-525 describe + gsift.py:800-871 _descr_flat for m keypoints (float64 sample
 * geometry, float64 bincount in the reference's corner-major order, float32 normalisation) */
void fso_sift_describe(const fso_sift *s, const float *kx, const float *ky, const float *ksize,
                       const float *kangle, const int32_t *koct, const int32_t *klayer,
                       const int64_t *sel, int64_t m, float *desc) {
    const int HS = (D_WIDTH + 2) * (D_WIDTH + 2) * (D_BINS + 2);
    double *hist = malloc(sizeof(double) * HS);
    for (int64_t q = 0; q < m; q++) {
        int64_t k = sel[q];
        int o = koct[k], li = klayer[k];
        int R = s->R[o], C = s->C[o];
        double f = 1.0 / (ldexp(1.0, o) * 0.5);
        double px = (double)kx[k] * f, py = (double)ky[k] * f;
        double scl = (double)ksize[k] * f * 0.5;
        float ori = 360.0f - kangle[k];
        if (fabsf(ori - 360.0f) < 1e-4f) ori = 0.0f;
        int64_t rad = (int64_t)rint(((3.0 * scl) * 1.4142135623730951) * 5 * 0.5);
        int64_t rlim = (int64_t)sqrt((double)((int64_t)R * R + (int64_t)C * C));
        if (rad > rlim) rad = rlim;
        double hw = D_SCL * scl;
        float t = ori * (float)(3.14159265358979323846 / 180.0);
        double cos_t = (double)cos_f(t) / hw, sin_t = (double)sin_f(t) / hw;
        int64_t pr = (int64_t)rint(py), pc = (int64_t)rint(px);
        const float *M = s->MAG[o] + (size_t)(li - 1) * R * C, *A = s->ANG[o] + (size_t)(li - 1) * R * C;
        const float obs = (float)(D_BINS / 360.0);
        for (int b = 0; b < HS; b++) hist[b] = 0.0;
        for (int corner = 0; corner < 8; corner++) {
            int dr = corner >> 2, dc = (corner >> 1) & 1, dor = corner & 1;
            for (int64_t i = -rad; i <= rad; i++)
                for (int64_t j = -rad; j <= rad; j++) {
                    double oi = (double)(float)i, oj = (double)(float)j;
                    double c_rot = oj * cos_t - oi * sin_t;
                    double r_rot = oj * sin_t + oi * cos_t;
                    double rbin = (r_rot + 2.0) - 0.5;
                    double cbin = (c_rot + 2.0) - 0.5;
                    int64_t yy = pr + i, xx = pc + j;
                    if (!(rbin > -1 && rbin < D_WIDTH && cbin > -1 && cbin < D_WIDTH && yy > 0 && yy < R - 1 &&
                          xx > 0 && xx < C - 1))
                        continue;
                    double w = exp((c_rot * c_rot + r_rot * r_rot) * (-1.0 / (D_WIDTH * D_WIDTH * 0.5)));
                    size_t g = (size_t)yy * C + xx;
                    double mg = (double)M[g] * w;
                    float obin = (A[g] - ori) * obs;
                    double r0 = floor(rbin), c0 = floor(cbin);
                    float o0f = floorf(obin);
                    double rb = rbin - r0, cb = cbin - c0;
                    float ob = obin - o0f;
                    int o0 = (int)o0f;
                    if (o0 < 0) o0 += D_BINS;
                    if (o0 >= D_BINS) o0 -= D_BINS;
                    double v_r1 = mg * rb, v_r0 = mg - v_r1;
                    double v_rc11 = v_r1 * cb, v_rc10 = v_r1 - v_rc11;
                    double v_rc01 = v_r0 * cb, v_rc00 = v_r0 - v_rc01;
                    double v111 = v_rc11 * ob, v110 = v_rc11 - v111;
                    double v101 = v_rc10 * ob, v100 = v_rc10 - v101;
                    double v011 = v_rc01 * ob, v010 = v_rc01 - v011;
                    double v001 = v_rc00 * ob, v000 = v_rc00 - v001;
                    double vals[8] = {v000, v001, v010, v011, v100, v101, v110, v111};
                    int base = (((int)r0 + 1) * (D_WIDTH + 2) + (int)c0 + 1) * (D_BINS + 2) + o0;
                    int idx = base + dr * (D_WIDTH + 2) * (D_BINS + 2) + dc * (D_BINS + 2) + dor;
                    hist[idx] += vals[corner];
                }
        }
        float hf[(D_WIDTH + 2) * (D_WIDTH + 2) * (D_BINS + 2)];
        for (int b = 0; b < HS; b++) hf[b] = (float)hist[b];
        for (int rc = 0; rc < (D_WIDTH + 2) * (D_WIDTH + 2); rc++) {
            hf[rc * (D_BINS + 2) + 0] += hf[rc * (D_BINS + 2) + D_BINS];
            hf[rc * (D_BINS + 2) + 1] += hf[rc * (D_BINS + 2) + D_BINS + 1];
        }
        float dst[128], sq[128];
        for (int rr = 0; rr < D_WIDTH; rr++)
            for (int cc = 0; cc < D_WIDTH; cc++)
                for (int ob = 0; ob < D_BINS; ob++)
                    dst[(rr * D_WIDTH + cc) * D_BINS + ob] = hf[((rr + 1) * (D_WIDTH + 2) + cc + 1) * (D_BINS + 2) + ob];
        for (int b = 0; b < 128; b++) sq[b] = dst[b] * dst[b];
        float nrm = sqrtf(np_sum_f(sq, 128));
        float thr = nrm * 0.2f;
        for (int b = 0; b < 128; b++) dst[b] = dst[b] < thr ? dst[b] : thr;
        for (int b = 0; b < 128; b++) sq[b] = dst[b] * dst[b];
        nrm = sqrtf(np_sum_f(sq, 128));
        float sc = 512.0f / (nrm > 1e-7f ? nrm : 1e-7f);
        for (int b = 0; b < 128; b++) desc[q * 128 + b] = dst[b] * sc;
    }
    free(hist);
}

/* features.py:80-82 RootSIFT uint8 encoding: L1-normalise, sqrt, *512, clip 0..255, truncate */
void fso_rootsift_u8(const float *desc, int64_t m, uint8_t *out) {
    for (int64_t q = 0; q < m; q++) {
        float a[128];
        for (int b = 0; b < 128; b++) a[b] = fabsf(desc[q * 128 + b]);
        float l1 = np_sum_f(a, 128) + 1e-7f;
        for (int b = 0; b < 128; b++) {
            float v = sqrtf(desc[q * 128 + b] / l1) * 512.0f;
            v = v < 0.f ? 0.f : (v > 255.f ? 255.f : v);
            out[q * 128 + b] = (uint8_t)v;
        }
    }
}

/* out[i] = fma(a[i], b[i], c[i]): a * b + c with one rounding (C99 fma, correctly rounded), the
 * fused multiply-add of the device's float64 inlier tests (verify.hip), which numpy lacks */
void fso_fma(const double *a, const double *b, const double *c, double *out, int64_t n) {
    for (int64_t i = 0; i < n; i++) out[i] = fma(a[i], b[i], c[i]);
}
