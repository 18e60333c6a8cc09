// Learned match filter on gfx950 (no counterpart in the reference): a 264-256-256-1 MLP with
// ReLU and a sigmoid output, BF16 weights / activations, FP32 accumulation on MFMA, trained
// with binary cross-entropy and Adam by the kernels below.
//
//   k_gemm<TA, TB, EPI>: C[M][N] = op(A)[M][K] op(B)[K][N], BF16 operands staged through LDS,
//       32x32x16 BF16 MFMA, 64x64 tiles per 256-thread block (2x2 waves of 32x32), K in steps
//       of 32.  Epilogues: bias + ReLU -> BF16 activations; plain FP32 store (weight
//       gradients); ReLU mask of a saved activation -> BF16 (backward through ReLU).
//   k_head: the 256 -> 1 layer, sigmoid, loss and its gradient, and dH2 = dz3 w3 masked by
//       ReLU, one wave per row; per-block partial sums for dw3 / db3 / loss.
//   k_colsum: bias gradients (column sums of dZ) and the head partials.
//   k_adam: Adam on FP32 master weights, refreshing the BF16 copies.
#include "fsx_common.h"
#include <math.h>

namespace fsx {

typedef __bf16 mbf16x8_t __attribute__((ext_vector_type(8)));
typedef float mf32x16_t __attribute__((ext_vector_type(16)));

__device__ __forceinline__ unsigned short m_f2bf(float x) {
    const unsigned b = __float_as_uint(x);
    return (unsigned short)((b + 0x7fffu + ((b >> 16) & 1u)) >> 16);
}
__device__ __forceinline__ float m_bf2f(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

constexpr int EPI_BIAS_RELU = 0;    // C16 = bf16(relu(acc + bias[n]))
constexpr int EPI_F32 = 1;          // C32 = acc
constexpr int EPI_MASK = 2;         // C16 = bf16(acc * (mask[m][n] > 0)), mask: bf16 [M][ldc]

constexpr int GT = 64;              // tile edge
constexpr int GK = 32;              // K step
constexpr int GKP = GK + 8;         // padded LDS row (bf16 elements)

// op(A)[m][k] = TA ? A[k * lda + m] : A[m * lda + k]; op(B)[k][n] = TB ? B[n * ldb + k] : B[k * ldb + n]
template <bool TA, bool TB, int EPI>
__global__ __launch_bounds__(256) void k_gemm(const unsigned short* __restrict__ A, int lda,
                                              const unsigned short* __restrict__ Bm, int ldb, int M, int N, int K,
                                              float* __restrict__ C32, unsigned short* __restrict__ C16, int ldc,
                                              const float* __restrict__ bias, const unsigned short* __restrict__ mask,
                                              int kchunk) {
    __shared__ __attribute__((aligned(16))) unsigned short sA[GT][GKP];   // [m][k]
    __shared__ __attribute__((aligned(16))) unsigned short sB[GT][GKP];   // [n][k]
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m0 = blockIdx.y * GT, n0 = blockIdx.x * GT;
    const int wm = (wv >> 1) * 32, wn = (wv & 1) * 32;
    // split K: block z sums k in [z * kchunk, (z + 1) * kchunk) into the z-th M x ldc slice of C32
    const int kb = blockIdx.z * kchunk, ke = min(K, kb + kchunk);
    if (EPI == EPI_F32) C32 += (size_t)blockIdx.z * M * ldc;
    mf32x16_t acc = {};
    for (int k0 = kb; k0 < ke; k0 += GK) {
        // stage A tile: 64 (m) x 32 (k)
        for (int e = tid; e < GT * GK; e += 256) {
            int m, k;
            if (TA) {            // contiguous along m
                k = e / GT;
                m = e - k * GT;
            } else {             // contiguous along k
                m = e / GK;
                k = e - m * GK;
            }
            const int gm = m0 + m, gk = k0 + k;
            unsigned short v = 0;
            if (gm < M && gk < ke) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
            sA[m][k] = v;
        }
        // stage B tile: 32 (k) x 64 (n), stored [n][k]
        for (int e = tid; e < GT * GK; e += 256) {
            int n, k;
            if (TB) {            // contiguous along k
                n = e / GK;
                k = e - n * GK;
            } else {             // contiguous along n
                k = e / GT;
                n = e - k * GT;
            }
            const int gn = n0 + n, gk = k0 + k;
            unsigned short v = 0;
            if (gn < N && gk < ke) v = TB ? Bm[(size_t)gn * ldb + gk] : Bm[(size_t)gk * ldb + gn];
            sB[n][k] = v;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < GK; kk += 16) {
            const uint4 av = *(const uint4*)&sA[wm + (lane & 31)][kk + (lane >> 5) * 8];
            const uint4 bv = *(const uint4*)&sB[wn + (lane & 31)][kk + (lane >> 5) * 8];
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mbf16x8_t, av),
                                                          __builtin_bit_cast(mbf16x8_t, bv), acc, 0, 0, 0);
        }
        __syncthreads();
    }
    // acc[i*4+j] -> row wm + i*8 + (lane/32)*4 + j, column wn + lane % 32
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int m = m0 + wm + i * 8 + (lane >> 5) * 4 + j, n = n0 + wn + (lane & 31);
            if (m >= M || n >= N) continue;
            const float a = acc[i * 4 + j];
            const size_t o = (size_t)m * ldc + n;
            if (EPI == EPI_BIAS_RELU) C16[o] = m_f2bf(fmaxf(a + bias[n], 0.0f));
            else if (EPI == EPI_F32) C32[o] = a;
            else if (EPI == EPI_MASK) C16[o] = m_f2bf(m_bf2f(mask[o]) > 0.0f ? a : 0.0f);
        }
}

// C[m][n] = sum over the split-K partials part[z][m][n], z = 0 .. splits-1 in order
__global__ __launch_bounds__(256) void k_splitk_sum(const float* __restrict__ part, int splits, int M, int N, int ldc,
                                                    float* __restrict__ C, int ldo) {
    const long long e = (long long)blockIdx.x * 256 + threadIdx.x;
    if (e >= (long long)M * N) return;
    const int m = (int)(e / N), n = (int)(e - (long long)m * N);
    float a = 0.0f;
    for (int z = 0; z < splits; z++) a += part[((size_t)z * M + m) * ldc + n];
    C[(size_t)m * ldo + n] = a;
}

// Head: z = H2[b] . w3 + b3, p = sigmoid(z); with labels: loss = BCE, dz = (p - y) / nb,
// dZ2[b][j] = dz * w3[j] * (H2[b][j] > 0).  One wave per row, 4 rows per block; block
// partials part[blk][0..257] = (sum_b dz H2[b][j] for j < 256, sum dz, sum loss).
constexpr int HID = 256;
constexpr int N_INP_MLP = 288;      // input width (264 features padded to the GEMM's K step)

__global__ __launch_bounds__(256) void k_head(const unsigned short* __restrict__ H2, const float* __restrict__ w3,
                                              const float* __restrict__ b3, const float* __restrict__ y, int nb,
                                              float* __restrict__ p_out, unsigned short* __restrict__ dZ2,
                                              float* __restrict__ part) {
    __shared__ float s_part[4][HID + 2];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int b = blockIdx.x * 4 + wv;
    float h[4], w[4];
    float z = 0.0f;
    if (b < nb) {
#pragma unroll
        for (int u = 0; u < 4; u++) {
            h[u] = m_bf2f(H2[(size_t)b * HID + lane * 4 + u]);
            w[u] = w3[lane * 4 + u];
            z += h[u] * w[u];
        }
    } else {
#pragma unroll
        for (int u = 0; u < 4; u++) h[u] = w[u] = 0.0f;
    }
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) z += __shfl_xor(z, off);
    z += b3[0];
    const float p = 1.0f / (1.0f + expf(-z));
    float dz = 0.0f, loss = 0.0f;
    if (b < nb) {
        if (lane == 0 && p_out) p_out[b] = p;
        if (y) {
            const float t = y[b];
            // BCE with logits, stable: max(z,0) - z t + log(1 + exp(-|z|))
            loss = fmaxf(z, 0.0f) - z * t + log1pf(expf(-fabsf(z)));
            dz = (p - t) / (float)nb;
#pragma unroll
            for (int u = 0; u < 4; u++)
                dZ2[(size_t)b * HID + lane * 4 + u] = m_f2bf(h[u] > 0.0f ? dz * w[u] : 0.0f);
        }
    }
    if (!part) return;
#pragma unroll
    for (int u = 0; u < 4; u++) s_part[wv][lane * 4 + u] = dz * h[u];
    if (lane == 0) {
        s_part[wv][HID] = dz;
        s_part[wv][HID + 1] = loss;
    }
    __syncthreads();
    for (int j = tid; j < HID + 2; j += 256)
        part[(size_t)blockIdx.x * (HID + 2) + j] = (s_part[0][j] + s_part[1][j]) + (s_part[2][j] + s_part[3][j]);
}

// part[chunk][j] = sum of rows [chunk * CS_ROWS, ...) of in[r * ld + j] (float32 or bf16 input):
// a fixed order, so the column sums are the same on every run
constexpr int CS_ROWS = 256;

template <bool BF16>
__global__ __launch_bounds__(256) void k_colsum(const void* __restrict__ in, int nr, int ld, int ncol,
                                                float* __restrict__ part) {
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j >= ncol) return;
    const int r0 = blockIdx.y * CS_ROWS, r1 = min(nr, r0 + CS_ROWS);
    float s = 0.0f;
    for (int r = r0; r < r1; r++)
        s += BF16 ? m_bf2f(((const unsigned short*)in)[(size_t)r * ld + j]) : ((const float*)in)[(size_t)r * ld + j];
    part[(size_t)blockIdx.y * ncol + j] = s;
}

// Adam (bias-corrected): m = b1 m + (1 - b1) g; v = b2 v + (1 - b2) g^2;
// w -= lr * (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + eps); refreshes the bf16 copy
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ w, float* __restrict__ m, float* __restrict__ v,
                                              const float* __restrict__ g, unsigned short* __restrict__ w16, int n,
                                              float lr, float beta1, float beta2, float eps, float c1, float c2) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / c1) / (sqrtf(vi / c2) + eps);
    w[i] = wi;
    if (w16) w16[i] = m_f2bf(wi);
}

__global__ __launch_bounds__(256) void k_to_bf16(const float* __restrict__ in, unsigned short* __restrict__ out, int n) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = m_f2bf(in[i]);
}

}  // namespace fsx

using namespace fsx;

// ---- fused inference -------------------------------------------------------------------
// p = sigmoid(relu(relu(X W1 + b1) W2 + b2) . w3 + b3) for rows of X (bf16 [n][288]) in one
// launch, the hidden activations rounded to bf16 as the training forward rounds them.  Each
// wave owns 32 rows and computes the transposed products (hidden units as the MFMA's M, rows
// as N), so a lane's accumulators are 16 hidden units of its own row: the 32x32x16 BF16 MFMA's
// C layout holds hidden rows 8 (j / 4) + 4 (lane / 32) + j % 4 of a 32-unit tile, which after
// bias, ReLU and rounding are exactly the lane's B operand for two k-steps of layer 2 when
// that layer's weights take the same order (mi_k2).  The weights are packed once
// (k_mlp_pack) in the operands' lane order, one 16-byte piece per lane, k-step and tile, in
// 16 KB stages: 9 stages of layer 1 (2 k-steps each, all 8 hidden tiles, so a wave keeps 8
// independent accumulators), then 8 of layer 2 (one 32-unit output tile, all 16 k-steps).  A
// block stages them one at a time in LDS (double-buffered) for its 8 waves, so each weight
// byte read from L2 serves 256 rows.
constexpr int MI_K1 = N_INP_MLP / 16;            // 18 k-steps of layer 1
constexpr int MI_K2 = HID / 16;                  // 16 k-steps of layer 2
constexpr int MI_NT = HID / 32;                  // 8 hidden tiles per layer
constexpr int MI_S1 = MI_K1 / 2;                 // 9 layer-1 stages of 2 k-steps
constexpr int MI_STAGE = 8192;                   // bf16 per stage (16 KB)
constexpr int MI_NSTAGE = MI_S1 + MI_NT;         // 17
constexpr int MI_PACK = MI_NSTAGE * MI_STAGE;    // bf16 in the packed weights
constexpr int MI_WAVES = 8, MI_ROWS = 32 * MI_WAVES;
static_assert(2 * MI_NT * 64 * 8 == MI_STAGE && MI_K2 * 64 * 8 == MI_STAGE && 2 * MI_S1 == MI_K1, "stage layout");

// hidden unit of k-step element i for lane half h in the layer-2 order (see above)
__host__ __device__ __forceinline__ int mi_k2(int s, int h, int i) { return 16 * s + 8 * (i >> 2) + 4 * h + (i & 3); }

__global__ __launch_bounds__(256) void k_mlp_pack(const unsigned short* __restrict__ W1, const unsigned short* __restrict__ W2,
                                                  unsigned short* __restrict__ pk) {
    const int e = blockIdx.x * 256 + threadIdx.x;
    if (e >= MI_PACK) return;
    const int g = e / MI_STAGE, r = e % MI_STAGE;
    const int i = r & 7, l = (r >> 3) & 63;
    if (g < MI_S1) {                                        // [k-step 2 g + sl][tile t][lane][i]
        const int t = (r >> 9) & 7, sl = r >> 12;
        pk[e] = W1[(16 * (2 * g + sl) + 8 * (l >> 5) + i) * HID + 32 * t + (l & 31)];
    } else {                                                // tile g - 9: [k-step s][lane][i]
        const int s = r >> 9;
        pk[e] = W2[mi_k2(s, l >> 5, i) * HID + 32 * (g - MI_S1) + (l & 31)];
    }
}

__global__ __launch_bounds__(64 * MI_WAVES) void k_mlp_infer(const unsigned short* __restrict__ X, int n,
                                                            const unsigned short* __restrict__ pk,
                                                            const float* __restrict__ b1, const float* __restrict__ b2,
                                                            const float* __restrict__ w3, const float* __restrict__ b3,
                                                            float* __restrict__ p_out) {
    __shared__ __attribute__((aligned(16))) unsigned short s_w[2][MI_STAGE];
    constexpr int NTH = 64 * MI_WAVES;
    static_assert(2 * NTH * 8 == MI_STAGE, "stage copy: two 16-byte pieces per thread");
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, h = lane >> 5;
    const long long row = (long long)blockIdx.x * MI_ROWS + wv * 32 + (lane & 31);
    const bool live = row < n;
    const uint4* xr = (const uint4*)(X + (size_t)(live ? row : 0) * N_INP_MLP);
    // staging of the next stage: NP = 2 pieces per thread, kept in st0 / st1
    uint4 st0 = ((const uint4*)pk)[tid], st1 = ((const uint4*)pk)[tid + NTH];
    ((uint4*)s_w[0])[tid] = st0;
    ((uint4*)s_w[0])[tid + NTH] = st1;
    // layer-1 B operands of stage c: the lane's row, inputs 16 (2 c + sl) + 8 h .. + 7 (a lane
    // past the last row reads row 0 and writes nothing)
    uint4 x0 = xr[h], x1 = xr[2 + h];
    __syncthreads();
    mf32x16_t acc[MI_NT];
#pragma unroll
    for (int t = 0; t < MI_NT; t++)
#pragma unroll
        for (int j = 0; j < 16; j++) acc[t][j] = 0.0f;
    for (int c = 0; c < MI_S1; c++) {
        const uint4* nsrc = (const uint4*)(pk + (size_t)(c + 1) * MI_STAGE);   // stage c + 1 < 17 always exists
        st0 = nsrc[tid];
        st1 = nsrc[tid + NTH];
        const int cn = c + 1 < MI_S1 ? c + 1 : c;
        const uint4 y0 = xr[4 * cn + h], y1 = xr[4 * cn + 2 + h];
        const uint4* w = (const uint4*)s_w[c & 1];
#pragma unroll
        for (int t = 0; t < MI_NT; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mbf16x8_t, w[t * 64 + lane]),
                                                             __builtin_bit_cast(mbf16x8_t, x0), acc[t], 0, 0, 0);
#pragma unroll
        for (int t = 0; t < MI_NT; t++)
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(__builtin_bit_cast(mbf16x8_t, w[(MI_NT + t) * 64 + lane]),
                                                             __builtin_bit_cast(mbf16x8_t, x1), acc[t], 0, 0, 0);
        // the other buffer was last read before the previous barrier
        ((uint4*)s_w[(c + 1) & 1])[tid] = st0;
        ((uint4*)s_w[(c + 1) & 1])[tid + NTH] = st1;
        x0 = y0;
        x1 = y1;
        __syncthreads();
    }
    // bias, ReLU, bf16: tile t's elements j = 4q .. 4q + 3 -> k-step 2 t + q / 2, pieces 2 (q % 2), + 1
    unsigned hb[MI_K2][4];
#pragma unroll
    for (int t = 0; t < MI_NT; t++)
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 bb = *(const float4*)(b1 + 32 * t + 8 * q + 4 * h);
            const unsigned short r0 = m_f2bf(fmaxf(acc[t][4 * q] + bb.x, 0.0f));
            const unsigned short r1 = m_f2bf(fmaxf(acc[t][4 * q + 1] + bb.y, 0.0f));
            const unsigned short r2 = m_f2bf(fmaxf(acc[t][4 * q + 2] + bb.z, 0.0f));
            const unsigned short r3 = m_f2bf(fmaxf(acc[t][4 * q + 3] + bb.w, 0.0f));
            hb[2 * t + (q >> 1)][2 * (q & 1)] = (unsigned)r0 | ((unsigned)r1 << 16);
            hb[2 * t + (q >> 1)][2 * (q & 1) + 1] = (unsigned)r2 | ((unsigned)r3 << 16);
        }
    float z = 0.0f;
    for (int t = 0; t < MI_NT; t++) {
        const int g = MI_S1 + t;
        if (g + 1 < MI_NSTAGE) {
            const uint4* nsrc = (const uint4*)(pk + (size_t)(g + 1) * MI_STAGE);
            st0 = nsrc[tid];
            st1 = nsrc[tid + NTH];
        }
        const uint4* w = (const uint4*)s_w[g & 1];
        mf32x16_t a2;
#pragma unroll
        for (int j = 0; j < 16; j++) a2[j] = 0.0f;
#pragma unroll
        for (int s = 0; s < MI_K2; s++)
            a2 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                __builtin_bit_cast(mbf16x8_t, w[s * 64 + lane]),
                __builtin_bit_cast(mbf16x8_t, make_uint4(hb[s][0], hb[s][1], hb[s][2], hb[s][3])), a2, 0, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const float4 bb = *(const float4*)(b2 + 32 * t + 8 * q + 4 * h);
            const float4 ww = *(const float4*)(w3 + 32 * t + 8 * q + 4 * h);
            z += m_bf2f(m_f2bf(fmaxf(a2[4 * q] + bb.x, 0.0f))) * ww.x;
            z += m_bf2f(m_f2bf(fmaxf(a2[4 * q + 1] + bb.y, 0.0f))) * ww.y;
            z += m_bf2f(m_f2bf(fmaxf(a2[4 * q + 2] + bb.z, 0.0f))) * ww.z;
            z += m_bf2f(m_f2bf(fmaxf(a2[4 * q + 3] + bb.w, 0.0f))) * ww.w;
        }
        if (g + 1 < MI_NSTAGE) {
            ((uint4*)s_w[(g + 1) & 1])[tid] = st0;
            ((uint4*)s_w[(g + 1) & 1])[tid + NTH] = st1;
        }
        __syncthreads();
    }
    z += __shfl_xor(z, 32);
    z += b3[0];
    if (live && h == 0) p_out[row] = 1.0f / (1.0f + expf(-z));
}

extern "C" {

// C = op(A) op(B) with the chosen epilogue; operands bf16 (uint16 bit patterns), row-major with
// leading dimensions lda / ldb / ldc (elements).  epi 0: C16 = relu(acc + bias); 1: C32 = acc;
// 2: C16 = acc where mask > 0 else 0.  K must be a multiple of 32.
int fsx_mlp_gemm(int ta, int tb, int epi, const void* A, int lda, const void* Bm, int ldb, int M, int N, int K,
                 float* C32, void* C16, int ldc, const float* bias, const void* mask, void* stream) {
    FSX_REQUIRE(A && Bm && M > 0 && N > 0 && K > 0 && (K % GK) == 0, "bad GEMM shape M=%d N=%d K=%d", M, N, K);
    FSX_REQUIRE(epi >= 0 && epi <= 2, "bad epilogue %d", epi);
    FSX_REQUIRE(epi != 0 || (C16 && bias), "epilogue 0 needs C16 and bias");
    FSX_REQUIRE(epi != 1 || C32, "epilogue 1 needs C32");
    FSX_REQUIRE(epi != 2 || (C16 && mask), "epilogue 2 needs C16 and mask");
    FSX_REQUIRE(lda > 0 && ldb > 0 && ldc >= N, "bad leading dimensions");
    const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT);
    hipStream_t st = (hipStream_t)stream;
    const unsigned short* a = (const unsigned short*)A;
    const unsigned short* b = (const unsigned short*)Bm;
    unsigned short* c16 = (unsigned short*)C16;
    const unsigned short* mk = (const unsigned short*)mask;
#define FSX_G(TA_, TB_, E_)                                                                          \
    if (!!ta == TA_ && !!tb == TB_ && epi == E_)                                                     \
        hipLaunchKernelGGL((k_gemm<TA_, TB_, E_>), grid, dim3(256), 0, st, a, lda, b, ldb, M, N, K, C32, c16, ldc, \
                           bias, mk, K);
    FSX_G(false, false, 0) FSX_G(false, false, 1) FSX_G(false, false, 2)
    FSX_G(true, false, 0) FSX_G(true, false, 1) FSX_G(true, false, 2)
    FSX_G(false, true, 0) FSX_G(false, true, 1) FSX_G(false, true, 2)
    FSX_G(true, true, 0) FSX_G(true, true, 1) FSX_G(true, true, 2)
#undef FSX_G
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Scratch bytes of fsx_mlp_gemm_splitk for an M x N product in `splits` K chunks.
int fsx_mlp_splitk_scratch(int M, int N, int splits, size_t* bytes) {
    FSX_REQUIRE(M > 0 && N > 0 && splits >= 1 && splits <= 256 && bytes, "bad split-K arguments");
    *bytes = (size_t)splits * M * N * sizeof(float);
    return FSX_OK;
}

// C32 (M x N, leading dimension ldc) = op(A) op(B) with K split into `splits` chunks of whole
// K steps computed by separate blocks into scratch, then summed chunk by chunk in order
// (deterministic).  For products with a small M x N and a long K (weight gradients over a
// batch), which a single pass would give to a handful of blocks.
int fsx_mlp_gemm_splitk(int ta, int tb, const void* A, int lda, const void* Bm, int ldb, int M, int N, int K, float* C32,
                        int ldc, float* scratch, int splits, void* stream) {
    FSX_REQUIRE(A && Bm && C32 && scratch && M > 0 && N > 0 && K > 0 && (K % GK) == 0,
                "bad split-K GEMM shape M=%d N=%d K=%d", M, N, K);
    FSX_REQUIRE(splits >= 1 && splits <= 256 && lda > 0 && ldb > 0 && ldc >= N, "bad split-K arguments");
    const int kchunk = ((K / GK + splits - 1) / splits) * GK;
    const int nz = (K + kchunk - 1) / kchunk;
    const dim3 grid((N + GT - 1) / GT, (M + GT - 1) / GT, nz);
    hipStream_t st = (hipStream_t)stream;
    const unsigned short* a = (const unsigned short*)A;
    const unsigned short* b = (const unsigned short*)Bm;
#define FSX_G(TA_, TB_)                                                                                        \
    if (!!ta == TA_ && !!tb == TB_)                                                                            \
        hipLaunchKernelGGL((k_gemm<TA_, TB_, EPI_F32>), grid, dim3(256), 0, st, a, lda, b, ldb, M, N, K, scratch, \
                           nullptr, N, nullptr, nullptr, kchunk);
    FSX_G(false, false) FSX_G(true, false) FSX_G(false, true) FSX_G(true, true)
#undef FSX_G
    hipLaunchKernelGGL(k_splitk_sum, dim3((unsigned)(((long long)M * N + 255) / 256)), dim3(256), 0, st, scratch, nz,
                       M, N, N, C32, ldc);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Head layer over nb rows of H2 (bf16, 256 wide): p_out (may be NULL) = sigmoid(H2 w3 + b3);
// with labels y (may be NULL for inference): dZ2 (bf16, nb x 256) and block partials part
// (ceil(nb / 4) x 258 floats: sum dz H2[:, j], sum dz, sum loss).
int fsx_mlp_head(const void* H2, const float* w3, const float* b3, const float* y, int nb, float* p_out, void* dZ2,
                 float* part, void* stream) {
    FSX_REQUIRE(H2 && w3 && b3 && nb > 0, "bad head arguments");
    FSX_REQUIRE(!y || (dZ2 && part), "training head needs dZ2 and partials");
    hipLaunchKernelGGL(k_head, dim3((nb + 3) / 4), dim3(256), 0, (hipStream_t)stream, (const unsigned short*)H2, w3,
                       b3, y, nb, p_out, (unsigned short*)dZ2, part);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Bytes of scratch fsx_mlp_colsum needs for nr rows of ncol columns.
int fsx_mlp_colsum_scratch(int nr, int ncol, size_t* bytes) {
    FSX_REQUIRE(bytes && nr >= 0 && ncol > 0, "bad colsum arguments");
    const int nch = (nr + CS_ROWS - 1) / CS_ROWS;
    *bytes = (size_t)4 * ncol * (nch + (nch + CS_ROWS - 1) / CS_ROWS + 1);
    return FSX_OK;
}

// out[j] = sum_r in[r][j] for j < ncol (in: bf16 if bf16 != 0, else float32; row stride ld), by
// fixed chunks of rows (deterministic); scratch: fsx_mlp_colsum_scratch bytes of device memory
int fsx_mlp_colsum(const void* in, int bf16, int nr, int ld, int ncol, float* out, float* scratch, void* stream) {
    FSX_REQUIRE(in && out && scratch && nr > 0 && ncol > 0 && ld >= ncol, "bad colsum arguments");
    hipStream_t st = (hipStream_t)stream;
    int nch = (nr + CS_ROWS - 1) / CS_ROWS;
    const dim3 grid((ncol + 255) / 256, nch);
    if (bf16) hipLaunchKernelGGL(k_colsum<true>, grid, dim3(256), 0, st, in, nr, ld, ncol, nch == 1 ? out : scratch);
    else hipLaunchKernelGGL(k_colsum<false>, grid, dim3(256), 0, st, in, nr, ld, ncol, nch == 1 ? out : scratch);
    FSX_LAUNCH_CHECK();
    // reduce the chunk sums (float32) until one row remains
    float* src = scratch;
    float* dst = scratch + (size_t)nch * ncol;
    while (nch > 1) {
        const int n2 = (nch + CS_ROWS - 1) / CS_ROWS;
        hipLaunchKernelGGL(k_colsum<false>, dim3((ncol + 255) / 256, n2), dim3(256), 0, st, (const void*)src, nch,
                           ncol, ncol, n2 == 1 ? out : dst);
        FSX_LAUNCH_CHECK();
        nch = n2;
        float* t = src;
        src = dst;
        dst = t;
    }
    return FSX_OK;
}

int fsx_mlp_adam(float* w, float* m, float* v, const float* g, void* w16, int n, float lr, float beta1, float beta2,
                 float eps, int step, void* stream) {
    FSX_REQUIRE(w && m && v && g && n > 0 && step >= 1, "bad Adam arguments");
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_adam, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, w, m, v, g,
                       (unsigned short*)w16, n, lr, beta1, beta2, eps, c1, c2);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// Bytes of the packed inference weights (fsx_mlp_infer_pack).
int fsx_mlp_infer_pack_bytes(size_t* bytes) {
    FSX_REQUIRE(bytes, "null pointer argument");
    *bytes = (size_t)MI_PACK * 2;
    return FSX_OK;
}

// Packs the bf16 weights W1 ([288][256]) and W2 ([256][256]) for fsx_mlp_infer into `packed`
// (fsx_mlp_infer_pack_bytes bytes, 16-byte aligned); run again after the weights change.
int fsx_mlp_infer_pack(const void* W1, const void* W2, void* packed, void* stream) {
    FSX_REQUIRE(W1 && W2 && packed, "null pointer argument");
    FSX_REQUIRE(((uintptr_t)packed & 15) == 0, "packed weights must be 16-byte aligned");
    hipLaunchKernelGGL(k_mlp_pack, dim3((MI_PACK + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       (const unsigned short*)W1, (const unsigned short*)W2, (unsigned short*)packed);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

// p_out[r] = sigmoid(relu(relu(X[r] W1 + b1) W2 + b2) . w3 + b3) for r < n, X bf16 [n][288]
// (16-byte aligned rows), weights packed by fsx_mlp_infer_pack; hidden activations rounded to
// bf16 as in the training forward (fsx_mlp_gemm epilogue 0 + fsx_mlp_head).
int fsx_mlp_infer(const void* X, int n, const void* packed, const float* b1, const float* b2, const float* w3,
                  const float* b3, float* p_out, void* stream) {
    FSX_REQUIRE(n >= 0, "bad row count");
    if (n == 0) return FSX_OK;
    FSX_REQUIRE(X && packed && b1 && b2 && w3 && b3 && p_out, "null pointer argument");
    FSX_REQUIRE(((uintptr_t)X & 15) == 0 && ((uintptr_t)packed & 15) == 0 && ((uintptr_t)b1 & 15) == 0 &&
                    ((uintptr_t)b2 & 15) == 0 && ((uintptr_t)w3 & 15) == 0,
                "X, packed weights, b1, b2 and w3 must be 16-byte aligned");
    hipLaunchKernelGGL(k_mlp_infer, dim3((n + MI_ROWS - 1) / MI_ROWS), dim3(64 * MI_WAVES), 0, (hipStream_t)stream,
                       (const unsigned short*)X, n, (const unsigned short*)packed, b1, b2, w3, b3, p_out);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

int fsx_mlp_to_bf16(const float* in, void* out, int n, void* stream) {
    FSX_REQUIRE(in && out && n > 0, "bad conversion arguments");
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, in,
                       (unsigned short*)out, n);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
