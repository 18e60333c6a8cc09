// Brute-force ratio matching of float descriptors (matching.py:86-100 _ratio_match_bf with
// matching.py:103-110 _unique) on gfx950, exact against the FP64 match-set definition:
// d(q, t) = sum_k (a_k - b_k)^2 accumulated in float64 in k order; best = argmin d (ties: lower
// index); b2 = the second smallest d; accept iff sqrt(b1) < ratio * sqrt(b2) in float64; per
// target the smallest b1 wins, ties to the lower query; output sorted by target.
//
// Pipeline (one launch each, all pairs of a batch at once):
//   quantize: per descriptor its float64 norm, a low-precision copy (BF16, or FP8 e4m3 with a
//             per-descriptor scale) and the exact norm of its quantization error;
//   screen:   MFMA over every (query, target) of a pair with FP32 accumulation (BF16
//             32x32x16, or FP8 32x32x64), keeping the K smallest approximate distances per query;
//   rescore:  the K candidates in float64 with the definition's arithmetic, then a certificate:
//             any target outside the K has approximate distance >= D_K (the K-th), hence true
//             distance >= D_K - eps_q, where eps_q bounds the screening error of query q
//             (quantization errors via |a^.b^ - a.b| <= |a| |b - b^| + |b^| |a - a^|, plus the
//             FP32 accumulation and rounding terms).  Accept / reject decisions that hold for
//             every consistent true value are final; the rest go to
//   fallback: one block per uncertain query, exact float64 over all targets;
//   unique / compact: per target the smallest (b1, q), then (q, t) lists in target order.
#include "fsx_common.h"
#include <float.h>
#include <math.h>
#include <string.h>
#include <algorithm>
#include <vector>

namespace fsx {

constexpr int FD = 128;            // descriptor length
constexpr int FK_MAX = 8;          // candidates kept per query

struct FSet {
    long long off;     // first descriptor in the concatenated arrays
    int n;
    int pad;
};

struct FPair {
    int a, b;          // queries from set a, targets from set b
    long long q_off;   // first global query (prefix of n_a)
    long long t_off;   // first per-target slot (prefix of n_b)
    long long o_off;   // first output slot (prefix of min(n_a, n_b))
};

struct FBatch {
    const float* desc;            // [Ntot][128] float32
    const FSet* set;
    const FPair* pair;
    int n_set, n_pair;
    long long n_kp, n_query, n_targets;
    double ratio;
    int prec;                     // 1 = BF16, 2 = FP8 e4m3
    int K;                        // candidates kept (2..8)
    // per descriptor
    unsigned short* qbf;          // [Ntot][128] bf16 bits (prec 1)
    uint8_t* q8;                  // [Ntot][128] e4m3 codes (prec 2)
    float* scale;                 // [Ntot] fp8 scale, a power of two (1 for bf16)
    uint8_t* sexp;                // [Ntot] its E8M0 code (exponent + 127)
    double* n2;                   // [Ntot] |a|^2 (float64, k order)
    float* nrm;                   // [Ntot] |a| rounded up
    float* qerr;                  // [Ntot] |a - dequant(a^)| rounded up
    unsigned* set_nmax;           // [n_set] max |b| (float bits)
    unsigned* set_emax;           // [n_set] max quantization error (float bits)
    // per query
    float* topd;                  // [n_query][FK_MAX] screened distances, ascending
    int* topi;                    // [n_query][FK_MAX] their target indices
    int* best;                    // accepted target, -1 none
    double* b1;                   // its exact distance
    unsigned long long* tkey;     // [n_targets] min accepted b1 bits
    int* tq;                      // [n_targets] min query among equal b1
    int* fb_list;                 // [n_query] uncertain queries (global index)
    int* counters;                // [0] fallback count, [1] max |d~ - d| (float bits),
                                  // [2..3] sum |d~ - d| (double), [4..5] compared count (u64)
    double* stat_part;            // [n_stat][3] per rescore block: max, sum, count
    long long n_stat;
    int* out_q;
    int* out_t;
    int* out_n;
};

__device__ __forceinline__ int fpair_of(const FBatch& B, long long g) {
    int lo = 0, hi = B.n_pair - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.pair[mid].q_off <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int fset_of(const FBatch& B, long long k) {
    int lo = 0, hi = B.n_set - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (B.set[mid].off <= k) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// float32 -> bfloat16, round to nearest even (finite inputs)
__device__ __forceinline__ unsigned short f2bf(float x) {
    const unsigned b = __float_as_uint(x);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// |v| <= 448 -> OCP FP8 e4m3 code (bias 7, subnormals m * 2^-9), round to nearest even
__device__ __forceinline__ unsigned f2e4m3(float v) {
    const unsigned sign = v < 0.0f ? 0x80u : 0u;
    const float a = fabsf(v);
    if (a == 0.0f) return sign;
    const unsigned b = __float_as_uint(a);
    int e = (int)(b >> 23) - 127;
    unsigned code;
    if (e < -6) {
        code = (unsigned)rintf(a * 512.0f);          // 0..8; 8 is the smallest normal's code
    } else {
        const unsigned mant = b & 0x7fffffu;
        unsigned q = mant >> 20;
        const unsigned rem = mant & 0xfffffu;
        if (rem > 0x80000u || (rem == 0x80000u && (q & 1u))) q++;
        if (q == 8u) {
            q = 0u;
            e++;
        }
        code = ((unsigned)(e + 7) << 3) | q;
        code = code > 0x7eu ? 0x7eu : code;          // 448
    }
    return sign | code;
}

// the value of an e4m3 code as a float (exact: at most 4 significant bits, exponents -9..8)
__device__ __forceinline__ float e4m3_float(unsigned c) {
    const unsigned ex = (c >> 3) & 15u, m = c & 7u;
    const float mag = ex == 0 ? (float)m * 0x1p-9f : __uint_as_float(((ex + 120u) << 23) | (m << 20));
    return (c & 0x80u) ? -mag : mag;
}

// ------------------------------------------------------------------------------------------
// Quantize + resets

// 32 lanes per descriptor: lane j holds elements 4j .. 4j+3 (coalesced 16-byte loads, 8- or
// 4-byte stores); norms, maxima and error sums reduce across the 32 lanes
__device__ __forceinline__ double red32_add(double v) {
#pragma unroll
    for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 32);
    return v;
}

__global__ __launch_bounds__(256) void k_f_quant(const FBatch B) {
    const long long tid = (long long)blockIdx.x * 256 + threadIdx.x;
    const long long stride = (long long)gridDim.x * 256;
    for (long long t = tid; t < B.n_targets; t += stride) {
        B.tkey[t] = ~0ull;
        B.tq[t] = 0x7fffffff;
    }
    if (tid < 8) B.counters[tid] = 0;
    const int j = threadIdx.x & 31;
    for (long long k = tid >> 5; k < B.n_kp; k += stride >> 5) {
        const float4 v = ((const float4*)(B.desc + k * FD))[j];
        const float x[4] = {v.x, v.y, v.z, v.w};
        double n2 = 0.0;
        float amax = 0.0f;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            n2 += (double)x[u] * (double)x[u];
            amax = fmaxf(amax, fabsf(x[u]));
        }
        n2 = red32_add(n2);
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) amax = fmaxf(amax, __shfl_xor(amax, off, 32));
        double e2 = 0.0;
        float sc = 1.0f;
        if (B.prec == 1) {
            unsigned short h[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                h[u] = f2bf(x[u]);
                const double d = (double)x[u] - (double)bf2f(h[u]);
                e2 += d * d;
            }
            ((uint2*)(B.qbf + k * FD))[j] = make_uint2((unsigned)h[0] | ((unsigned)h[1] << 16),
                                                       (unsigned)h[2] | ((unsigned)h[3] << 16));
        } else {
            // the smallest power of two with amax / sc <= 448 (exact divisions below), kept
            // within E8M0's range
            int ex = 0;
            if (amax > 0.0f) {
                const float m = frexpf(amax / 448.0f, &ex);     // amax / 448 = m 2^ex, m in [0.5, 1)
                if (m == 0.5f) ex -= 1;
                ex = ex < -127 ? -127 : (ex > 127 ? 127 : ex);
            }
            sc = ldexpf(1.0f, ex);
            // x / sc as x * 2^-ex: both are the correctly rounded x 2^-ex (2^-ex is a float
            // for |ex| <= 127), the multiply without a divide sequence
            const float isc = ldexpf(1.0f, -ex);
            // e4m3 codes by the hardware conversion (round to nearest even; the same code as
            // f2e4m3 for every float in [-448, 448] up to the sign of zero: fsx_fp8_cvt_probe
            // checks all of them)
            float y[4];
#pragma unroll
            for (int u = 0; u < 4; u++) y[u] = fminf(fmaxf(x[u] * isc, -448.0f), 448.0f);
            unsigned w = __builtin_amdgcn_cvt_pk_fp8_f32(y[0], y[1], 0, false);
            w = __builtin_amdgcn_cvt_pk_fp8_f32(y[2], y[3], w, true);
#pragma unroll
            for (int u = 0; u < 4; u++) {
                const double d = (double)x[u] - (double)sc * (double)e4m3_float((w >> (8 * u)) & 0xffu);
                e2 += d * d;
            }
            ((unsigned*)(B.q8 + k * FD))[j] = w;
        }
        e2 = red32_add(e2);
        if (j == 0) {
            B.scale[k] = sc;
            B.sexp[k] = (uint8_t)(B.prec == 2 ? (int)ilogbf(sc) + 127 : 127);
            B.n2[k] = n2;
            // upper bounds as float32 (the float64 sums' rounding is far below the margin)
            const float nr = __double2float_ru(sqrt(n2) * (1.0 + 1e-12));
            const float er = __double2float_ru(sqrt(e2) * (1.0 + 1e-12));
            B.nrm[k] = nr;
            B.qerr[k] = er;
        }
    }
}

// per set: max |b| and max quantization error (one block per set, no contended atomics)
__global__ __launch_bounds__(256) void k_f_setmax(const FBatch B) {
    __shared__ float s_n[4], s_e[4];
    const FSet S = B.set[blockIdx.x];
    float mn = 0.0f, me = 0.0f;
    for (int i = threadIdx.x; i < S.n; i += 256) {
        mn = fmaxf(mn, B.nrm[S.off + i]);
        me = fmaxf(me, B.qerr[S.off + i]);
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        mn = fmaxf(mn, __shfl_xor(mn, off));
        me = fmaxf(me, __shfl_xor(me, off));
    }
    if ((threadIdx.x & 63) == 0) {
        s_n[threadIdx.x >> 6] = mn;
        s_e[threadIdx.x >> 6] = me;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        B.set_nmax[blockIdx.x] = __float_as_uint(fmaxf(fmaxf(s_n[0], s_n[1]), fmaxf(s_n[2], s_n[3])));
        B.set_emax[blockIdx.x] = __float_as_uint(fmaxf(fmaxf(s_e[0], s_e[1]), fmaxf(s_e[2], s_e[3])));
    }
}

// ------------------------------------------------------------------------------------------
// Screen: block = 4 waves x 32 queries; targets stream through LDS in tiles of 64.

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef int i32x8_t __attribute__((ext_vector_type(8)));

// Candidate lists hold keys: a key is a non-negative float whose low `ib` bits are replaced by
// the target's index (ib = bits of the target count), so one float compare orders by screened
// distance truncated to the key's mantissa bits, ties by index, and one v_med3_f32 per slot
// inserts (no separate index array, no selects).  The truncated value (index bits cleared) is
// <= the screened distance, so the K-th key's value bounds every target outside the list from
// below: such a target's key is larger, hence its truncated distance, and with it its screened
// distance, is at least that value.  An empty slot holds +inf; real keys are finite (the
// distance is finite and its exponent field is kept).
__device__ __forceinline__ float make_key(float d, unsigned himask, unsigned idx) {
    return __uint_as_float((__float_as_uint(fmaxf(d, 0.0f)) & himask) | idx);
}

// Insertion into an ascending key list: the new entry j is the median of (kl[j-1], key,
// kl[j]) (entry 0: the smaller of key and kl[0]), written from the top down in place.  A key
// above kl[K-1] (+inf included) changes nothing.
template <int K>
__device__ __forceinline__ void key_insert(float key, float* kl) {
#pragma unroll
    for (int j = K - 1; j >= 1; j--) kl[j] = __builtin_amdgcn_fmed3f(kl[j - 1], key, kl[j]);
    // entry 0: both are non-negative floats or +inf, ordered as unsigned integers
    kl[0] = __uint_as_float(min(__float_as_uint(kl[0]), __float_as_uint(key)));
}

// Score gate of a query with squared norm qn2 whose K-th screened distance is at least tau: a
// target whose accumulator (a.b - |b|^2 / 2) does not exceed it has screened distance
// qn2 - 2 acc >= tau (the margin covers the float32 roundings of both sides).
__device__ __forceinline__ float gate_of(float qn2, float tau) {
    return 0.5f * (qn2 - tau) - 0x1p-20f * (fabsf(qn2) + fabsf(tau)) - 1e-30f;
}

// the gate of a key list: -FLT_MAX while it has an empty slot (every real target enters;
// rows past the end carry -inf and never do)
__device__ __forceinline__ float gate_of_list(float qn2, float kK, unsigned himask) {
    return kK == INFINITY ? -FLT_MAX : gate_of(qn2, __uint_as_float(__float_as_uint(kK) & himask));
}

template <int PREC, int K, int FS_TILE, int QB>                   // QB: query blocks per wave
__global__ __launch_bounds__(256) void k_f_screen(const FBatch B) {
    constexpr int ROWB = PREC == 1 ? FD * 2 : FD;                 // bytes per staged row
    // LDS row pitch: 16 bytes more than a row, so the 16-byte operand reads of 16 consecutive
    // lanes (consecutive rows) fall on 16 disjoint 4-bank groups
    constexpr int PITCH = ROWB + 16;
    constexpr int TBYTES = FS_TILE * PITCH;                       // one staged tile
    constexpr int NV = FS_TILE * ROWB / 16 / 256;                 // 16-byte pieces per thread per tile
    static_assert(NV * 16 * 256 == FS_TILE * ROWB && FS_TILE <= 256, "tile staging layout");
    constexpr int NBUF = 2;                                       // next tile fetched during this one
    constexpr int SBYTES = NBUF * TBYTES > 256 * QB * K * 4 ? NBUF * TBYTES : 256 * QB * K * 4;
    __shared__ __attribute__((aligned(16))) uint8_t s_rows[SBYTES];
    __shared__ __attribute__((aligned(16))) float s_c0[NBUF][FS_TILE];   // -|b|^2 / 2, -inf past the end
    __shared__ int s_sx[NBUF][FS_TILE];                                   // FP8: E8M0 scale of the row
    // FP8 stages each lane's 16 accumulator values in LDS for its insertion loop; BF16 inserts
    // straight from the registers, element by element (its MFMA time per step is twice FP8's,
    // so the per-element branches cost relatively less, and the LDS it saves buys a fourth wave)
    constexpr bool STAGED = PREC == 2;
    __shared__ __attribute__((aligned(16))) float s_acc[STAGED ? 256 * 16 : 4];
    float* s_mk = (float*)s_rows;
    const FPair P = B.pair[blockIdx.y];
    const FSet sa = B.set[P.a], sb = B.set[P.b];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if ((int)blockIdx.x * 128 * QB >= sa.n) return;                 // block-uniform
    const bool pair_ok = sa.n >= 2 && sb.n >= 2;
    // key layout of this pair: target indices below 2^ib
    const int ib = 32 - __clz(max(sb.n, 1));
    const unsigned himask = ~((1u << ib) - 1u);
    // query operands, norms and key lists of the QB blocks
    bf16x8_t qb[QB][8];
    i32x8_t q8[QB][2];
    int qsx[QB];
    float qn2[QB], hthr[QB];
    bool valid[QB];
    float kl[QB][K];
#pragma unroll
    for (int u = 0; u < QB; u++) {
        const int q = blockIdx.x * 128 * QB + (wv * QB + u) * 32 + (lane & 31);
        valid[u] = q < sa.n && pair_ok;
        const long long ka = sa.off + (q < sa.n ? q : 0);
        if (PREC == 1) {
            const uint4* src = (const uint4*)(B.qbf + ka * FD);
#pragma unroll
            for (int s = 0; s < 8; s++) qb[u][s] = __builtin_bit_cast(bf16x8_t, src[2 * s + (lane >> 5)]);
        } else {
            const uint4* src = (const uint4*)(B.q8 + ka * FD);
#pragma unroll
            for (int s = 0; s < 2; s++) {
                const uint4 v0 = src[4 * s + 2 * (lane >> 5)], v1 = src[4 * s + 2 * (lane >> 5) + 1];
                q8[u][s] = i32x8_t{(int)v0.x, (int)v0.y, (int)v0.z, (int)v0.w, (int)v1.x, (int)v1.y, (int)v1.z, (int)v1.w};
            }
            qsx[u] = (int)B.sexp[ka];
        }
        qn2[u] = (float)B.n2[ka];
        // an invalid query gates at +inf: hthr - acc is never negative, nothing enters
        hthr[u] = valid[u] ? -FLT_MAX : INFINITY;
#pragma unroll
        for (int j = 0; j < K; j++) kl[u][j] = INFINITY;
    }
    uint4 pv[NV];
    float pc0 = 0.0f;
    int psx = 127;
    auto fetch = [&](int base) {   // 16-byte loads of tile `base` into registers, all threads
        const int nt = min(FS_TILE, sb.n - base);
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int u = tid + 256 * k;
            const int r = u / (ROWB / 16), c = u - r * (ROWB / 16);
            pv[k] = make_uint4(0, 0, 0, 0);
            if (r < nt) {
                const uint8_t* row = PREC == 1 ? (const uint8_t*)(B.qbf + (sb.off + base + r) * FD)
                                               : B.q8 + (sb.off + base + r) * FD;
                pv[k] = ((const uint4*)row)[c];
            }
        }
        if (tid < FS_TILE) {
            const bool ok = tid < nt;
            pc0 = ok ? -0.5f * (float)B.n2[sb.off + base + tid] : -INFINITY;
            psx = ok && PREC == 2 ? (int)B.sexp[sb.off + base + tid] : 127;
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int u = tid + 256 * k;
            const int r = u / (ROWB / 16), c = u - r * (ROWB / 16);
            *(uint4*)(s_rows + buf * TBYTES + r * PITCH + 16 * c) = pv[k];
        }
        if (tid < FS_TILE) {
            s_c0[buf][tid] = pc0;
            s_sx[buf][tid] = psx;
        }
    };
    fetch(0);
    store(0);
    __syncthreads();
    int buf = 0;
    for (int base = 0; base < sb.n; base += FS_TILE) {
        const bool more = base + FS_TILE < sb.n;   // block-uniform
        if (more) fetch(base + FS_TILE);           // in flight during this tile's MFMA work
        const uint8_t* tile = s_rows + buf * TBYTES;
#pragma unroll
        for (int h = 0; h < FS_TILE / 32; h++) {
            // C operand: -|b|^2 / 2 of the lane's 16 accumulator rows i*8 + (lane/32)*4 + j
            f32x16_t c0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float4 v = *(const float4*)&s_c0[buf][h * 32 + i * 8 + (lane >> 5) * 4];
                c0[4 * i] = v.x;
                c0[4 * i + 1] = v.y;
                c0[4 * i + 2] = v.z;
                c0[4 * i + 3] = v.w;
            }
            f32x16_t acc[QB];
#pragma unroll
            for (int u = 0; u < QB; u++) acc[u] = c0;
            const int rl = h * 32 + (lane & 31);
            const uint4* rp = (const uint4*)(tile + rl * PITCH);
            if (PREC == 1) {
#pragma unroll
                for (int s = 0; s < 8; s++) {
                    const bf16x8_t a = __builtin_bit_cast(bf16x8_t, rp[2 * s + (lane >> 5)]);
#pragma unroll
                    for (int u = 0; u < QB; u++) acc[u] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, qb[u][s], acc[u], 0, 0, 0);
                }
            } else {
                const int asx = s_sx[buf][rl];
#pragma unroll
                for (int s = 0; s < 2; s++) {
                    const uint4 v0 = rp[4 * s + 2 * (lane >> 5)], v1 = rp[4 * s + 2 * (lane >> 5) + 1];
                    const i32x8_t a = {(int)v0.x, (int)v0.y, (int)v0.z, (int)v0.w,
                                       (int)v1.x, (int)v1.y, (int)v1.z, (int)v1.w};
                    // e4m3 operands (format 0) with the rows' / queries' E8M0 scales (byte 0)
#pragma unroll
                    for (int u = 0; u < QB; u++)
                        acc[u] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, q8[u][s], acc[u], 0, 0, 0, asx, 0,
                                                                                 qsx[u]);
                }
            }
            // row of element e: (e / 4) * 8 + (lane / 32) * 4 + e % 4 = rb + e + (e & 12)
            const unsigned rb = (unsigned)(base + h * 32 + (lane >> 5) * 4);
#pragma unroll
            for (int u = 0; u < QB; u++) {
                bool ins = false;                      // wave-uniform: some lane inserted
                if (STAGED) {
                    // pass mask of the lane's 16 values: bit e = acc[e] > hthr, the sign of
                    // hthr - acc[e] (never -0 or NaN: hthr is finite or +inf, acc finite or
                    // -inf), shifted in by one funnel shift per element
                    unsigned m = 0;
#pragma unroll
                    for (int e = 15; e >= 0; e--)
                        m = __builtin_amdgcn_alignbit(m, __float_as_uint(hthr[u] - acc[u][e]), 31);
                    if (__any(m != 0u)) {
                        ins = true;
                        if (m) {
                            // the lane's 16 values to its LDS slot (float4 chunks swizzled by
                            // lane & 3 against bank conflicts: element e at e ^ 4 (lane & 3)), then
                            // one iteration per passing element, software-pipelined: the next
                            // element's value is read while the current one is inserted
                            float* my = s_acc + ((wv * 64 + lane) << 4);
                            const int sw4 = (lane & 3) << 2;
#pragma unroll
                            for (int c = 0; c < 4; c++)
                                *(float4*)(my + ((4 * c) ^ sw4)) =
                                    make_float4(acc[u][4 * c], acc[u][4 * c + 1], acc[u][4 * c + 2], acc[u][4 * c + 3]);
                            int e = __builtin_ctz(m);
                            m &= m - 1;
                            float v = my[e ^ sw4];
                            while (true) {
                                const bool more_e = m != 0;
                                const int en = more_e ? __builtin_ctz(m) : e;
                                const float vn = my[en ^ sw4];
                                m &= m - 1;
                                // screened distance qn2 - 2v (2v is exact: one rounding)
                                const float d = __builtin_fmaf(-2.0f, v, qn2[u]);
                                key_insert<K>(make_key(d, himask, rb + (unsigned)(e + (e & 12))), kl[u]);
                                if (!more_e) break;
                                e = en;
                                v = vn;
                            }
                        }
                    }
                } else {
                    // one compare per element; an element passing in some lane is inserted
                    // from its accumulator register, an element passing in none costs the
                    // compare and a wave-uniform skip
#pragma unroll
                    for (int e = 0; e < 16; e++) {
                        const bool pass = acc[u][e] > hthr[u];
                        if (__any(pass)) {
                            ins = true;
                            if (pass) {
                                const float d = __builtin_fmaf(-2.0f, acc[u][e], qn2[u]);
                                key_insert<K>(make_key(d, himask, rb + (unsigned)(e + (e & 12))), kl[u]);
                            }
                        }
                    }
                }
                // both forms insert against the threshold of the step's start (a key above the
                // list's K-th leaves it unchanged) and refresh the gate once per step
                if (ins) {
                    hthr[u] = fmaxf(hthr[u], gate_of_list(qn2[u], kl[u][K - 1], himask));
                    // both half-waves of a query gate with the tighter threshold: an element of
                    // the union's top-K lies below both halves' K-th values.  The swap leaves
                    // the partner's value in one word and the lane's own in the other.
                    const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(hthr[u]),
                                                                     __float_as_uint(hthr[u]), false, false);
                    hthr[u] = fmaxf(__uint_as_float(sw[0]), __uint_as_float(sw[1]));
                }
            }
        }
        if (more) store(buf ^ 1);   // the other buffer was last read before the previous barrier
        __syncthreads();
        buf ^= 1;
    }
    // merge the two half-waves of each query
#pragma unroll
    for (int u = 0; u < QB; u++)
#pragma unroll
        for (int j = 0; j < K; j++) s_mk[(u * 256 + tid) * K + j] = kl[u][j];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < QB; u++) {
        if (lane < 32 && valid[u]) {
#pragma unroll
            for (int j = 0; j < K; j++) key_insert<K>(s_mk[(u * 256 + tid + 32) * K + j], kl[u]);
            const int q = blockIdx.x * 128 * QB + (wv * QB + u) * 32 + lane;
            const long long g = P.q_off + q;
#pragma unroll
            for (int j = 0; j < K; j++) {
                // value: the truncated screened distance (a lower bound of it); -1: empty slot
                const bool full = kl[u][j] != INFINITY;
                const unsigned kb = __float_as_uint(kl[u][j]);
                B.topd[g * FK_MAX + j] = full ? __uint_as_float(kb & himask) : INFINITY;
                B.topi[g * FK_MAX + j] = full ? (int)(kb & ~himask) : -1;
            }
        }
    }
}

// ------------------------------------------------------------------------------------------
// Rescore + certificate

// sum_k (a_k - b_k)^2 in float64, k in order (the definition's arithmetic)
__device__ __forceinline__ double dist64(const float* __restrict__ a, const float* __restrict__ b) {
    const float4* a4 = (const float4*)a;
    const float4* b4 = (const float4*)b;
    double acc = 0.0;
#pragma unroll 8
    for (int i = 0; i < FD / 4; i++) {
        const float4 x = a4[i], y = b4[i];
        double t = (double)x.x - (double)y.x;
        acc += t * t;
        t = (double)x.y - (double)y.y;
        acc += t * t;
        t = (double)x.z - (double)y.z;
        acc += t * t;
        t = (double)x.w - (double)y.w;
        acc += t * t;
    }
    return acc;
}

// bound on |d~ - d| for query k against any target of set s
__device__ __forceinline__ double screen_eps(const FBatch& B, long long ka, int s) {
    const double u = 0x1p-24;
    const double na = (double)B.nrm[ka], ea = (double)B.qerr[ka];
    const double nb = (double)__uint_as_float(B.set_nmax[s]), eb = (double)__uint_as_float(B.set_emax[s]);
    // FP32 accumulation of FD products plus the -|b|^2 / 2 start value, x2 margin
    const double gam = 2.0 * ((FD + 1) * u) / (1.0 - (FD + 1) * u);
    const double qa = na + ea, qb = nb + eb;
    // quantization: |a^.b^ - a.b| <= |a| |b - b^| + |b^| |a - a^|
    const double quant = na * eb + qb * ea;
    // accumulation over sum |a^_k b^_k| + |b|^2 / 2 <= qa qb + nb^2 / 2, then the float32
    // roundings of |a|^2, |b|^2 and of d~ = |a|^2 - 2 acc
    const double accum = gam * (qa * qb + 0.5 * nb * nb);
    const double rnd = 4.0 * u * (na * na + nb * nb + 2.0 * qa * qb);
    return (2.0 * (quant + accum) + rnd) * 1.01 + 1e-30;
}

__device__ __forceinline__ void f_accept(const FBatch& B, long long g, const FPair& P, int t, double b1) {
    B.best[g] = t;
    B.b1[g] = b1;
    atomicMin(&B.tkey[P.t_off + t], (unsigned long long)__double_as_longlong(b1));
}

// one 8-lane group per query: lane j rescores candidate j
__global__ __launch_bounds__(256) void k_f_rescore(const FBatch B) {
    const FPair P = B.pair[blockIdx.y];
    const FSet sa = B.set[P.a], sb = B.set[P.b];
    const int tid = threadIdx.x, sub = tid & 7;
    const int q = blockIdx.x * 32 + (tid >> 3);
    const bool inq = q < sa.n;
    const long long g = P.q_off + (inq ? q : 0);
    const bool live = inq && sa.n >= 2 && sb.n >= 2;
    if (inq && !live && sub == 0) B.best[g] = -1;
    const int K = B.K;
    if (K == 0) {                           // exact mode: every query takes the float64 pass
        if (live && sub == 0) {
            B.best[g] = -1;
            B.fb_list[atomicAdd(&B.counters[0], 1)] = (int)g;
        }
        return;
    }
    const long long ka = sa.off + (inq ? q : 0);
    const int kk = min(K, max(sb.n, 1));
    double dv = INFINITY;
    int iv = 0x7fffffff;
    float dt = INFINITY;
    float err = 0.0f;
    bool cand = false;
    if (live && sub < kk) {
        iv = B.topi[g * FK_MAX + sub];
        dt = B.topd[g * FK_MAX + sub];
        cand = iv >= 0 && iv < sb.n;  // a screened slot left empty (no finite distance) stays out
    }
    // Only candidates that can be among the two smallest exact distances are rescored.  The
    // exact distance lies within eps of the screened one, which lies in [dt, dt (1 + 2^(ib-23)))
    // (dt: the key's value, ib index bits truncated); U2, the second smallest upper bound of the
    // group's candidates, bounds the second smallest exact distance b2, so a candidate with
    // dt - eps >= U2 has distance >= b2 and changes neither b1, b2 nor the best index (a tie at
    // b1 = b2 fails the ratio test whatever the index).  Its row is not read.
    const double eps = live ? screen_eps(B, ka, P.b) : 0.0;
    const int ib = 32 - __clz(max(sb.n, 1));
    double m1 = cand ? (double)dt * (1.0 + ldexp(1.0, ib - 22)) + ldexp(1.0, ib - 149) + eps : INFINITY, m2 = INFINITY;
#pragma unroll
    for (int off = 1; off < 8; off <<= 1) {          // the two smallest over the 8 lanes of the group
        const double o1 = __shfl_xor(m1, off), o2 = __shfl_xor(m2, off);
        m2 = fmin(fmax(m1, o1), fmin(m2, o2));
        m1 = fmin(m1, o1);
    }
    const bool use = cand && (double)dt - eps < m2;
    if (use) {
        dv = dist64(B.desc + ka * FD, B.desc + (sb.off + iv) * FD);
        err = (float)fabs((double)dt - dv);
    } else {
        iv = 0x7fffffff;
    }
    // screening-error statistics: block partials (same-address atomics would serialise)
    {
        __shared__ double s_st[4][3];
        float emax = err;
        double esum = err;
        int ecnt = use ? 1 : 0;
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            emax = fmaxf(emax, __shfl_xor(emax, off));
            esum += __shfl_xor(esum, off);
            ecnt += __shfl_xor(ecnt, off);
        }
        if ((tid & 63) == 0) {
            s_st[tid >> 6][0] = emax;
            s_st[tid >> 6][1] = esum;
            s_st[tid >> 6][2] = ecnt;
        }
        __syncthreads();
        if (tid == 0) {
            double* o = B.stat_part + 3 * ((long long)blockIdx.y * gridDim.x + blockIdx.x);
            o[0] = fmax(fmax(s_st[0][0], s_st[1][0]), fmax(s_st[2][0], s_st[3][0]));
            o[1] = (s_st[0][1] + s_st[1][1]) + (s_st[2][1] + s_st[3][1]);
            o[2] = (s_st[0][2] + s_st[1][2]) + (s_st[2][2] + s_st[3][2]);
        }
    }
    // gather the group's exact values (the 8 lanes of a group are consecutive)
    double d[FK_MAX];
    int id[FK_MAX];
#pragma unroll
    for (int j = 0; j < FK_MAX; j++) {
        d[j] = __shfl(dv, (tid & ~7) + j);
        id[j] = __shfl(iv, (tid & ~7) + j);
    }
    const float dK = __shfl(dt, (tid & ~7) + (kk - 1));
    if (!live || sub != 0) return;
    // exact top-2 among the candidates
    double b1 = INFINITY, b2 = INFINITY;
    int i1 = 0x7fffffff;
#pragma unroll
    for (int j = 0; j < FK_MAX; j++) {
        if (j >= kk) continue;
        if (d[j] < b1 || (d[j] == b1 && id[j] < i1)) {
            b2 = b1;
            b1 = d[j];
            i1 = id[j];
        } else {
            b2 = fmin(b2, d[j]);
        }
    }
    if (sb.n <= kk) {                       // every target was rescored: exact
        if (sqrt(b1) < B.ratio * sqrt(b2)) f_accept(B, g, P, i1, b1);
        else B.best[g] = -1;
        return;
    }
    // targets outside the candidates: true distance >= L
    const double L = ((double)dK - eps) * (1.0 - 0x1p-40);
    const bool best_ok = b1 < L;
    const double lb2 = fmin(b2, L), lb1 = fmin(b1, L);
    if (best_ok && sqrt(b1) < B.ratio * sqrt(lb2)) {
        f_accept(B, g, P, i1, b1);
    } else if (sqrt(lb1) >= B.ratio * sqrt(b2)) {
        B.best[g] = -1;
    } else {
        B.best[g] = -1;
        const int slot = atomicAdd(&B.counters[0], 1);
        B.fb_list[slot] = (int)g;
    }
}

// exact float64 over all targets for the uncertain queries, one block each
__global__ __launch_bounds__(256) void k_f_fallback(const FBatch B) {
    __shared__ __attribute__((aligned(16))) float s_q[FD];
    __shared__ double s_b1[256], s_b2[256];
    __shared__ int s_i1[256];
    const int nfb = B.counters[0];
    for (int f = blockIdx.x; f < nfb; f += gridDim.x) {
        const long long g = B.fb_list[f];
        const FPair P = B.pair[fpair_of(B, g)];
        const FSet sa = B.set[P.a], sb = B.set[P.b];
        const long long ka = sa.off + (g - P.q_off);
        if (threadIdx.x < FD) s_q[threadIdx.x] = B.desc[ka * FD + threadIdx.x];
        __syncthreads();
        double b1 = INFINITY, b2 = INFINITY;
        int i1 = 0x7fffffff;
        for (int t = threadIdx.x; t < sb.n; t += 256) {
            const double d = dist64(s_q, B.desc + (sb.off + t) * FD);
            if (d < b1 || (d == b1 && t < i1)) {
                b2 = b1;
                b1 = d;
                i1 = t;
            } else {
                b2 = fmin(b2, d);
            }
        }
        s_b1[threadIdx.x] = b1;
        s_b2[threadIdx.x] = b2;
        s_i1[threadIdx.x] = i1;
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int u = 1; u < 256; u++) {
                const double c1 = s_b1[u], c2 = s_b2[u];
                const int ci = s_i1[u];
                if (c1 < b1 || (c1 == b1 && ci < i1)) {
                    b2 = fmin(b1, c2);
                    b1 = c1;
                    i1 = ci;
                } else {
                    b2 = fmin(b2, c1);
                }
            }
            if (sqrt(b1) < B.ratio * sqrt(b2)) f_accept(B, g, P, i1, b1);
        }
        __syncthreads();
    }
}

// statistics partials of the rescore blocks -> counters[1..5]
__global__ __launch_bounds__(1024) void k_f_stats(const FBatch B) {
    __shared__ double s_m[16], s_s[16], s_c[16];
    double m = 0.0, sm = 0.0, c = 0.0;
    for (long long i = threadIdx.x; i < B.n_stat; i += 1024) {
        m = fmax(m, B.stat_part[3 * i]);
        sm += B.stat_part[3 * i + 1];
        c += B.stat_part[3 * i + 2];
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) {
        m = fmax(m, __shfl_xor(m, off));
        sm += __shfl_xor(sm, off);
        c += __shfl_xor(c, off);
    }
    if ((threadIdx.x & 63) == 0) {
        s_m[threadIdx.x >> 6] = m;
        s_s[threadIdx.x >> 6] = sm;
        s_c[threadIdx.x >> 6] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 16; w++) {
            m = fmax(m, s_m[w]);
            sm += s_s[w];
            c += s_c[w];
        }
        const float mf = (float)m;
        B.counters[1] = (int)__float_as_uint(mf);
        *(double*)&B.counters[2] = sm;
        *(unsigned long long*)&B.counters[4] = (unsigned long long)c;
    }
}

__global__ __launch_bounds__(256) void k_f_unique2(const FBatch B) {
    const long long g = (long long)blockIdx.x * 256 + threadIdx.x;
    if (g >= B.n_query) return;
    const int t = B.best[g];
    if (t < 0) return;
    const FPair P = B.pair[fpair_of(B, g)];
    if ((unsigned long long)__double_as_longlong(B.b1[g]) == B.tkey[P.t_off + t])
        atomicMin(&B.tq[P.t_off + t], (int)(g - P.q_off));
}

__global__ __launch_bounds__(1024) void k_f_compact(const FBatch B) {
    __shared__ int s_w[16];
    const FPair P = B.pair[blockIdx.x];
    const int nb = B.set[P.b].n;
    const int* tq = B.tq + P.t_off;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int carry = 0;
    for (int base = 0; base < nb; base += 1024) {
        const int t = base + tid;
        const int v = (t < nb && tq[t] != 0x7fffffff) ? 1 : 0;
        int x = v;
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[wv] = x;
        __syncthreads();
        int pre = 0, tot = 0;
        for (int w2 = 0; w2 < 16; w2++) {
            pre += w2 < wv ? s_w[w2] : 0;
            tot += s_w[w2];
        }
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
// FP8 MFMA probe: A byte e of lane l (row l % 32, half h = l / 32) = e4m3 code of
// ((l%32)*7 + h*3 + e*5) % 11 - 5, B byte = code of ((l%32)*3 + h*5 + e*7) % 13 - 6 (small
// integers are exact in e4m3); out: 64 x 16 floats.
__global__ void k_fp8_probe(float* out) {
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    i32x8_t a, b;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        unsigned wa = 0, wb = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = 4 * w + u;
            wa |= f2e4m3((float)((r * 7 + h * 3 + e * 5) % 11 - 5)) << (8 * u);
            wb |= f2e4m3((float)((r * 3 + h * 5 + e * 7) % 13 - 6)) << (8 * u);
        }
        a[w] = (int)wa;
        b[w] = (int)wb;
    }
    f32x16_t acc = {};
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc, 0, 0, 0, 0x7f7f7f7f, 0, 0x7f7f7f7f);
#pragma unroll
    for (int i = 0; i < 16; i++) out[lane * 16 + i] = acc[i];
}

// The same operands with E8M0 scales per lane: lane l scales A row l % 32 by 2^((l%32) % 5 - 2)
// and B column l % 32 by 2^((l%32) % 3 - 1) (code = exponent + 127, byte 0), the form the
// screen uses for its power-of-two per-descriptor scales; out: 64 x 16 floats.
__global__ void k_fp8_scale_probe(float* out) {
    const int lane = threadIdx.x, r = lane & 31, h = lane >> 5;
    i32x8_t a, b;
#pragma unroll
    for (int w = 0; w < 8; w++) {
        unsigned wa = 0, wb = 0;
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int e = 4 * w + u;
            wa |= f2e4m3((float)((r * 7 + h * 3 + e * 5) % 11 - 5)) << (8 * u);
            wb |= f2e4m3((float)((r * 3 + h * 5 + e * 7) % 13 - 6)) << (8 * u);
        }
        a[w] = (int)wa;
        b[w] = (int)wb;
    }
    const int sa = r % 5 - 2 + 127, sb = r % 3 - 1 + 127;
    f32x16_t acc = {};
    acc = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(a, b, acc, 0, 0, 0, sa, 0, sb);
#pragma unroll
    for (int i = 0; i < 16; i++) out[lane * 16 + i] = acc[i];
}

// The hardware e4m3 conversion against f2e4m3 on every float in [-448, 448] (2^31 + 2 bit
// patterns: magnitudes 0 .. 448 with both signs; a +0 code for -0 counts as equal, both are
// zero): out[0] = mismatches, out[1] = the smallest mismatching bit pattern (all ones if none).
__global__ void k_fp8_cvt_init(unsigned long long* out) {
    out[0] = 0ull;
    out[1] = ~0ull;
}

__global__ __launch_bounds__(256) void k_fp8_cvt_probe(unsigned long long* out) {
    const unsigned long long n = 2ull * (0x43E00000ull + 1ull);
    unsigned long long bad = 0ull, first = ~0ull;
    for (unsigned long long i = (unsigned long long)blockIdx.x * 256 + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * 256) {
        const unsigned b = (unsigned)(i >> 1) | ((i & 1ull) ? 0x80000000u : 0u);
        const float x = __uint_as_float(b);
        const unsigned hw = __builtin_amdgcn_cvt_pk_fp8_f32(x, 0.0f, 0, false) & 0xffu;
        const unsigned sw = f2e4m3(x);
        if (hw != sw && ((hw | sw) & 0x7fu)) {         // +0 and -0 codes both count as zero
            bad++;
            first = first < b ? first : b;
        }
    }
    if (bad) {
        atomicAdd(&out[0], bad);
        atomicMin(&out[1], first);
    }
}

// ------------------------------------------------------------------------------------------
// Host side

static inline size_t fa256(size_t x) { return (x + 255) & ~(size_t)255; }

struct FPlan {
    std::vector<FSet> set;
    std::vector<FPair> pair;
    long long n_kp = 0, n_query = 0, n_targets = 0, n_out = 0, n_stat = 0;
    int max_na = 0;
    size_t off_stat, off_set, off_pair, off_qbf, off_q8, off_scale, off_sexp, off_n2, off_nrm, off_qerr, off_nmax, off_emax, off_topd,
        off_topi, off_best, off_b1, off_tkey, off_tq, off_fb, off_cnt, off_outq, off_outt, off_outn, total;
};

static int fplan(int n_set, const int* n_kp, int n_pair, const int* pa, const int* pb, FPlan& M) {
    FSX_REQUIRE(n_set >= 1 && n_pair >= 1 && n_kp && pa && pb, "need sets and pairs");
    M.set.resize(n_set);
    long long off = 0;
    for (int i = 0; i < n_set; i++) {
        FSX_REQUIRE(n_kp[i] >= 0, "bad set %d", i);
        M.set[i].off = off;
        M.set[i].n = n_kp[i];
        M.set[i].pad = 0;
        off += n_kp[i];
    }
    M.n_kp = off;
    M.pair.resize(n_pair);
    long long qo = 0, to = 0, oo = 0;
    for (int p = 0; p < n_pair; p++) {
        FSX_REQUIRE(pa[p] >= 0 && pa[p] < n_set && pb[p] >= 0 && pb[p] < n_set, "bad pair %d", p);
        FPair& P = M.pair[p];
        P.a = pa[p];
        P.b = pb[p];
        P.q_off = qo;
        P.t_off = to;
        P.o_off = oo;
        qo += M.set[P.a].n;
        to += M.set[P.b].n;
        oo += std::min(M.set[P.a].n, M.set[P.b].n);
        M.max_na = std::max(M.max_na, M.set[P.a].n);
    }
    M.n_query = qo;
    M.n_targets = to;
    M.n_out = oo;
    const size_t nk = (size_t)std::max(M.n_kp, 1LL), nq = (size_t)std::max(qo, 1LL), nt = (size_t)std::max(to, 1LL);
    size_t o = 0;
    M.off_set = o; o = fa256(o + sizeof(FSet) * n_set);
    M.off_pair = o; o = fa256(o + sizeof(FPair) * n_pair);
    M.off_qbf = o; o = fa256(o + 2 * FD * nk);
    M.off_q8 = o; o = fa256(o + FD * nk);
    M.off_scale = o; o = fa256(o + 4 * nk);
    M.off_sexp = o; o = fa256(o + nk);
    M.off_n2 = o; o = fa256(o + 8 * nk);
    M.off_nrm = o; o = fa256(o + 4 * nk);
    M.off_qerr = o; o = fa256(o + 4 * nk);
    M.off_nmax = o; o = fa256(o + 4 * (size_t)n_set);
    M.off_emax = o; o = fa256(o + 4 * (size_t)n_set);
    M.off_topd = o; o = fa256(o + 4 * FK_MAX * nq);
    M.off_topi = o; o = fa256(o + 4 * FK_MAX * nq);
    M.off_best = o; o = fa256(o + 4 * nq);
    M.off_b1 = o; o = fa256(o + 8 * nq);
    M.off_tkey = o; o = fa256(o + 8 * nt);
    M.off_tq = o; o = fa256(o + 4 * nt);
    M.off_fb = o; o = fa256(o + 4 * nq);
    M.off_cnt = o; o = fa256(o + 64);
    M.off_outq = o; o = fa256(o + 4 * (size_t)std::max(oo, 1LL));
    M.off_outt = o; o = fa256(o + 4 * (size_t)std::max(oo, 1LL));
    M.off_outn = o; o = fa256(o + 4 * (size_t)n_pair);
    M.n_stat = (long long)((M.max_na + 31) / 32) * n_pair;
    M.off_stat = o; o = fa256(o + 24 * (size_t)std::max(M.n_stat, 1LL));
    M.total = o;
    return FSX_OK;
}

}  // namespace fsx

using namespace fsx;

extern "C" {

int fsx_fmatch_workspace_bytes(int n_set, const int* n_kp, int n_pair, const int* pa, const int* pb, size_t* bytes,
                               long long* out_capacity) {
    FSX_REQUIRE(bytes, "null pointer argument");
    FPlan M;
    int rc = fplan(n_set, n_kp, n_pair, pa, pb, M);
    if (rc) return rc;
    *bytes = M.total;
    if (out_capacity) *out_capacity = M.n_out;
    return FSX_OK;
}

int fsx_fmatch_batch(int n_set, const int* n_kp, const float* desc_dev, int n_pair, const int* pa, const int* pb,
                     double ratio, int precision, int K, void* ws, size_t ws_bytes, long long* out_offsets,
                     void* stage_events_v, void* stream) {
    FSX_REQUIRE(desc_dev && ws, "null pointer argument");
    FSX_REQUIRE(precision == 1 || precision == 2, "precision must be 1 (BF16) or 2 (FP8 e4m3)");
    FSX_REQUIRE(K == 0 || K == 2 || K == 8, "K must be 0 (exact float64 for every query), 2 or 8");
    FPlan M;
    int rc = fplan(n_set, n_kp, n_pair, pa, pb, M);
    if (rc) return rc;
    FSX_REQUIRE(ws_bytes >= M.total, "fmatch workspace too small: %zu < %zu", ws_bytes, M.total);
    FSX_REQUIRE(n_pair <= 65535, "at most 65535 pairs per batch");
    hipStream_t st = (hipStream_t)stream;
    char* w = (char*)ws;
    FBatch b;
    memset(&b, 0, sizeof(b));
    b.desc = desc_dev;
    b.set = (const FSet*)(w + M.off_set);
    b.pair = (const FPair*)(w + M.off_pair);
    b.n_set = n_set;
    b.n_pair = n_pair;
    b.n_kp = M.n_kp;
    b.n_query = M.n_query;
    b.n_targets = M.n_targets;
    b.ratio = ratio;
    b.prec = precision;
    b.K = K;
    b.qbf = (unsigned short*)(w + M.off_qbf);
    b.q8 = (uint8_t*)(w + M.off_q8);
    b.scale = (float*)(w + M.off_scale);
    b.sexp = (uint8_t*)(w + M.off_sexp);
    b.n2 = (double*)(w + M.off_n2);
    b.nrm = (float*)(w + M.off_nrm);
    b.qerr = (float*)(w + M.off_qerr);
    b.set_nmax = (unsigned*)(w + M.off_nmax);
    b.set_emax = (unsigned*)(w + M.off_emax);
    b.topd = (float*)(w + M.off_topd);
    b.topi = (int*)(w + M.off_topi);
    b.best = (int*)(w + M.off_best);
    b.b1 = (double*)(w + M.off_b1);
    b.tkey = (unsigned long long*)(w + M.off_tkey);
    b.tq = (int*)(w + M.off_tq);
    b.fb_list = (int*)(w + M.off_fb);
    b.counters = (int*)(w + M.off_cnt);
    b.out_q = (int*)(w + M.off_outq);
    b.out_t = (int*)(w + M.off_outt);
    b.out_n = (int*)(w + M.off_outn);
    b.stat_part = (double*)(w + M.off_stat);
    b.n_stat = M.n_stat;
    rc = upload_async(w + M.off_set, M.set.data(), sizeof(FSet) * M.set.size(), st);
    if (rc) return rc;
    rc = upload_async(w + M.off_pair, M.pair.data(), sizeof(FPair) * M.pair.size(), st);
    if (rc) return rc;
    if (out_offsets)
        for (int p = 0; p < n_pair; p++) out_offsets[p] = M.pair[p].o_off;
    hipEvent_t* ev = (hipEvent_t*)stage_events_v;
    if (ev) FSX_HIP(hipEventRecord(ev[0], st));
    const long long nthr = std::max<long long>(std::max(32 * M.n_kp, M.n_targets), 1);
    hipLaunchKernelGGL(k_f_quant, dim3((unsigned)std::min<long long>((nthr + 255) / 256, 1 << 16)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_f_setmax, dim3((unsigned)n_set), dim3(256), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[1], st));
    if (M.n_query > 0 && K > 0) {
        const bool t128 = tune_get(TUNE_FSCREEN_TILE) == 128;
        const int qb = tune_get(TUNE_FSCREEN_QB);
        const dim3 grid((unsigned)((M.max_na + 128 * qb - 1) / (128 * qb)), (unsigned)n_pair);
#define FSX_SCREEN(P_, K_)                                                                            \
    if (precision == P_ && K == K_) {                                                                 \
        if (qb == 2) {                                                                                \
            if (t128) hipLaunchKernelGGL((k_f_screen<P_, K_, 128, 2>), grid, dim3(256), 0, st, b);    \
            else hipLaunchKernelGGL((k_f_screen<P_, K_, 64, 2>), grid, dim3(256), 0, st, b);          \
        } else {                                                                                      \
            if (t128) hipLaunchKernelGGL((k_f_screen<P_, K_, 128, 1>), grid, dim3(256), 0, st, b);    \
            else hipLaunchKernelGGL((k_f_screen<P_, K_, 64, 1>), grid, dim3(256), 0, st, b);          \
        }                                                                                             \
    }
        FSX_SCREEN(1, 8)
        FSX_SCREEN(1, 2)
        FSX_SCREEN(2, 8)
        FSX_SCREEN(2, 2)
#undef FSX_SCREEN
        FSX_LAUNCH_CHECK();
    }
    if (ev) FSX_HIP(hipEventRecord(ev[2], st));
    if (M.n_query > 0) {
        hipLaunchKernelGGL(k_f_rescore, dim3((unsigned)((M.max_na + 31) / 32), (unsigned)n_pair), dim3(256), 0, st, b);
        if (K > 0) hipLaunchKernelGGL(k_f_stats, dim3(1), dim3(1024), 0, st, b);
        hipLaunchKernelGGL(k_f_fallback, dim3(1024), dim3(256), 0, st, b);
        FSX_LAUNCH_CHECK();
    }
    if (ev) FSX_HIP(hipEventRecord(ev[3], st));
    if (M.n_query > 0)
        hipLaunchKernelGGL(k_f_unique2, dim3((unsigned)((M.n_query + 255) / 256)), dim3(256), 0, st, b);
    hipLaunchKernelGGL(k_f_compact, dim3((unsigned)n_pair), dim3(1024), 0, st, b);
    FSX_LAUNCH_CHECK();
    if (ev) FSX_HIP(hipEventRecord(ev[4], st));
    return FSX_OK;
}

// Results of the last fsx_fmatch_batch (synchronises): counts[n_pair], q/t at the pair offsets,
// stats[4] = fallback queries, max |d~ - d|, mean |d~ - d|, rescored candidates.
int fsx_fmatch_results(int n_set, const int* n_kp, int n_pair, const int* pa, const int* pb, const void* ws,
                       int* counts, int* q, int* t, double* stats, void* stream) {
    FSX_REQUIRE(ws && (counts || n_pair <= 0), "null pointer argument");
    FPlan M;
    int rc = fplan(n_set, n_kp, n_pair, pa, pb, M);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const char* w = (const char*)ws;
    int cnt[8];
    FSX_HIP(hipMemcpyAsync(counts, w + M.off_outn, 4 * (size_t)n_pair, hipMemcpyDeviceToHost, st));
    if (q) FSX_HIP(hipMemcpyAsync(q, w + M.off_outq, 4 * (size_t)M.n_out, hipMemcpyDeviceToHost, st));
    if (t) FSX_HIP(hipMemcpyAsync(t, w + M.off_outt, 4 * (size_t)M.n_out, hipMemcpyDeviceToHost, st));
    FSX_HIP(hipMemcpyAsync(cnt, w + M.off_cnt, sizeof(cnt), hipMemcpyDeviceToHost, st));
    FSX_HIP(hipStreamSynchronize(st));
    if (stats) {
        double sum;
        unsigned long long n;
        memcpy(&sum, cnt + 2, 8);
        memcpy(&n, cnt + 4, 8);
        float mx;
        memcpy(&mx, cnt + 1, 4);
        stats[0] = cnt[0];
        stats[1] = mx;
        stats[2] = n ? sum / (double)n : 0.0;
        stats[3] = (double)n;
    }
    return FSX_OK;
}

int fsx_fp8_scale_probe(float* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_fp8_scale_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_fp8_cvt_probe(unsigned long long* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_fp8_cvt_init, dim3(1), dim3(1), 0, (hipStream_t)stream, out_dev);
    hipLaunchKernelGGL(k_fp8_cvt_probe, dim3(2048), dim3(256), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_fp8_layout_probe(float* out_dev, void* stream) {
    FSX_REQUIRE(out_dev, "null pointer argument");
    hipLaunchKernelGGL(k_fp8_probe, dim3(1), dim3(64), 0, (hipStream_t)stream, out_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
