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

int fsx_mlp_to_bf16(const float* in, void* out, int n, void* stream) {
    FSX_REQUIRE(in && out && n > 0, "bad conversion arguments");
    hipLaunchKernelGGL(k_to_bf16, dim3((n + 255) / 256), dim3(256), 0, (hipStream_t)stream, in,
                       (unsigned short*)out, n);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
