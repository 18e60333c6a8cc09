// Contrast-limited adaptive histogram equalisation of uint8 photos (the input conditioning of
// features.py:60-71: cv2.createCLAHE(clipLimit=2.0, tileGridSize=(8, 8)).apply(gray)).
//
// OpenCV's 8-bit algorithm, step for step (float32 where it computes in float):
//   * tiles are cut from the image extended at the right / bottom by (tiles - size % tiles)
//     pixels (a whole tile when that side divides) with BORDER_REFLECT_101, unless both sides
//     divide;
//   * k_clahe_hist, one block per (tile, photo, row slab): 256-bin histogram (one LDS copy
//     per wave) added to the tile's global histogram; k_clahe_lut, one block per tile: bins
//     capped at `limit`, the excess spread as clipped / 256 per bin plus one each to
//     bins 0, s, 2s, ... (clipped % 256 of them, s = max(256 / that, 1)), inclusive scan,
//     LUT = round-half-even(float(cumsum) * (255 / tile area)) clamped to 0..255;
//   * k_clahe_apply, one thread per 1 or 4 pixels: bilinear blend of the four surrounding tiles' LUT
//     values with weights from x * (1 / tile width) - 0.5 (and the same along y), each float
//     operation rounded on its own (the library is built without contraction).
#include "fsx_common.h"

namespace {

constexpr int CL_THREADS = 256;

__global__ void k_clahe_zero(unsigned* hist, long long n) {
    for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x)
        hist[k] = 0;
}

// Histogram of one row slab of one tile (grid: tile, photo, slab), merged into the tile's
// global histogram with one atomic per non-empty bin.
__global__ void __launch_bounds__(CL_THREADS) k_clahe_hist(const uint8_t* src, int H, int W, int tiles_x, int th,
                                                           int tw, int slab_rows, unsigned* hist) {
    __shared__ unsigned s_hist[CL_THREADS / 64][256];
    const int tile = blockIdx.x, img = blockIdx.y;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < (CL_THREADS / 64) * 256; k += CL_THREADS) (&s_hist[0][0])[k] = 0;
    __syncthreads();
    const uint8_t* im = src + (size_t)img * H * W;
    const int r0 = blockIdx.z * slab_rows;
    const int r1 = r0 + slab_rows < th ? r0 + slab_rows : th;
    for (int r = r0 + wave; r < r1; r += CL_THREADS / 64) {
        int y = ty * th + r;
        if (y >= H) y = 2 * (H - 1) - y;                       // BORDER_REFLECT_101
        const uint8_t* row = im + (size_t)y * W;
        for (int cb = 0; cb < tw; cb += 8 * 64) {
            // up to 8 independent loads per lane in flight before the LDS atomics
            int val[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int c = cb + u * 64 + lane;
                int x = tx * tw + (c < tw ? c : 0);            // in-tile address for every lane
                if (x >= W) x = 2 * (W - 1) - x;
                val[u] = c < tw ? (int)row[x] : -1;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (val[u] >= 0) atomicAdd(&s_hist[wave][val[u]], 1u);
        }
    }
    __syncthreads();
    unsigned h = 0;
    for (int w = 0; w < CL_THREADS / 64; ++w) h += s_hist[w][threadIdx.x];
    if (h) atomicAdd(hist + ((size_t)img * gridDim.x + tile) * 256 + threadIdx.x, h);
}

// One block per (tile, photo): clip, redistribute, scan, LUT.
__global__ void __launch_bounds__(CL_THREADS) k_clahe_lut(const unsigned* hist, int limit, float scale, uint8_t* lut) {
    __shared__ int s_wave[CL_THREADS / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const size_t base = ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * 256;
    const int i = threadIdx.x;                                 // one bin per thread
    int h = (int)hist[base + i];
    if (limit > 0) {
        int excess = h > limit ? h - limit : 0;
        h -= excess;
        for (int d = 32; d >= 1; d >>= 1) excess += __shfl_xor(excess, d);
        if (lane == 0) s_wave[wave] = excess;
        __syncthreads();
        int clipped = 0;
        for (int w = 0; w < CL_THREADS / 64; ++w) clipped += s_wave[w];
        __syncthreads();
        const int batch = clipped / 256, residual = clipped - batch * 256;
        h += batch;
        if (residual) {
            const int step = 256 / residual > 1 ? 256 / residual : 1;
            if (i % step == 0 && i / step < residual) h += 1;
        }
    }
    int inc = h;                                               // inclusive scan over the 256 bins
    for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(inc, d);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_wave[wave] = inc;
    __syncthreads();
    for (int w = 0; w < wave; ++w) inc += s_wave[w];
    const float v = rintf((float)inc * scale);
    lut[base + i] = (uint8_t)(v < 0.f ? 0.f : (v > 255.f ? 255.f : v));
}

__device__ __forceinline__ uint8_t clahe_px(int v, int x, int tiles_x, float inv_tw, float ya, float ya1,
                                            const uint8_t* p1, const uint8_t* p2) {
    const float txf = (float)x * inv_tw - 0.5f;
    int tx1 = (int)floorf(txf);
    int tx2 = tx1 + 1;
    const float xa = txf - (float)tx1, xa1 = 1.0f - xa;
    tx1 = tx1 > 0 ? tx1 : 0;
    tx2 = tx2 < tiles_x - 1 ? tx2 : tiles_x - 1;
    const float res = ((float)p1[tx1 * 256 + v] * xa1 + (float)p1[tx2 * 256 + v] * xa) * ya1 +
                      ((float)p2[tx1 * 256 + v] * xa1 + (float)p2[tx2 * 256 + v] * xa) * ya;
    const float r = rintf(res);
    return (uint8_t)(r < 0.f ? 0.f : (r > 255.f ? 255.f : r));
}

// One thread per V consecutive pixels of a row (V = 4: 32-bit loads and stores, used when
// W % 4 == 0 and both buffers are 4-byte aligned).
template <int V>
__global__ void __launch_bounds__(CL_THREADS) k_clahe_apply(const uint8_t* src, int H, int W, int tiles_x, int tiles_y,
                                                            float inv_tw, float inv_th, const uint8_t* lut,
                                                            uint8_t* dst) {
    const int x0 = (blockIdx.x * CL_THREADS + threadIdx.x) * V;
    const int y = blockIdx.y, img = blockIdx.z;
    if (x0 >= W) return;
    const float tyf = (float)y * inv_th - 0.5f;
    int ty1 = (int)floorf(tyf);
    int ty2 = ty1 + 1;
    const float ya = tyf - (float)ty1, ya1 = 1.0f - ya;
    ty1 = ty1 > 0 ? ty1 : 0;
    ty2 = ty2 < tiles_y - 1 ? ty2 : tiles_y - 1;
    const uint8_t* L = lut + (size_t)img * tiles_x * tiles_y * 256;
    const uint8_t* p1 = L + (size_t)ty1 * tiles_x * 256;
    const uint8_t* p2 = L + (size_t)ty2 * tiles_x * 256;
    const size_t at = ((size_t)img * H + y) * W + x0;
    if (V == 4) {
        const unsigned v = *(const unsigned*)(src + at);
        unsigned o = 0;
        for (int k = 0; k < 4; ++k)
            o |= (unsigned)clahe_px((v >> (8 * k)) & 255, x0 + k, tiles_x, inv_tw, ya, ya1, p1, p2) << (8 * k);
        *(unsigned*)(dst + at) = o;
    } else {
        dst[at] = clahe_px(src[at], x0, tiles_x, inv_tw, ya, ya1, p1, p2);
    }
}

// LDS form of k_clahe_apply<4> for up to CL_LDS_TX tiles across: a block takes CL_ROWS rows of
// 1024 columns and first stages the LUTs of the tile rows those rows blend (at most 3 when
// CL_ROWS <= tile height: ty1 is monotone in y and moves by at most one over fewer than th
// rows), so the four gathers per pixel read LDS instead of going through the texture path.
constexpr int CL_ROWS = 16;
constexpr int CL_LDS_TX = 16;

__device__ __forceinline__ void clahe_tile_rows(int y, float inv_th, int tiles_y, int& ty1, int& ty2, float& ya,
                                                float& ya1) {
    const float tyf = (float)y * inv_th - 0.5f;
    ty1 = (int)floorf(tyf);
    ty2 = ty1 + 1;
    ya = tyf - (float)ty1;
    ya1 = 1.0f - ya;
    ty1 = ty1 > 0 ? ty1 : 0;
    ty2 = ty2 < tiles_y - 1 ? ty2 : tiles_y - 1;
}

__global__ void __launch_bounds__(CL_THREADS) k_clahe_apply_lds(const uint8_t* src, int H, int W, int tiles_x,
                                                                int tiles_y, float inv_tw, float inv_th,
                                                                const uint8_t* lut, uint8_t* dst) {
    __shared__ uint8_t s_lut[3 * CL_LDS_TX * 256];
    const int img = blockIdx.z;
    const int y0 = blockIdx.y * CL_ROWS;
    const int y1 = y0 + CL_ROWS < H ? y0 + CL_ROWS : H;
    int tyA, tyB, t_;
    float f_, g_;
    clahe_tile_rows(y0, inv_th, tiles_y, tyA, t_, f_, g_);
    clahe_tile_rows(y1 - 1, inv_th, tiles_y, t_, tyB, f_, g_);
    const int row_bytes = tiles_x * 256;
    const uint8_t* L = lut + ((size_t)img * tiles_y + tyA) * row_bytes;
    const int n_bytes = (tyB - tyA + 1) * row_bytes;            // <= 3 rows (checked by the host)
    for (int k = threadIdx.x * 4; k < n_bytes; k += CL_THREADS * 4) *(unsigned*)(s_lut + k) = *(const unsigned*)(L + k);
    __syncthreads();
    const int x0 = (blockIdx.x * CL_THREADS + threadIdx.x) * 4;
    if (x0 >= W) return;
    for (int y = y0; y < y1; ++y) {
        int ty1, ty2;
        float ya, ya1;
        clahe_tile_rows(y, inv_th, tiles_y, ty1, ty2, ya, ya1);
        const uint8_t* p1 = s_lut + (ty1 - tyA) * row_bytes;
        const uint8_t* p2 = s_lut + (ty2 - tyA) * row_bytes;
        const size_t at = ((size_t)img * H + y) * W + x0;
        const unsigned v = *(const unsigned*)(src + at);
        unsigned o = 0;
        for (int k = 0; k < 4; ++k)
            o |= (unsigned)clahe_px((v >> (8 * k)) & 255, x0 + k, tiles_x, inv_tw, ya, ya1, p1, p2) << (8 * k);
        *(unsigned*)(dst + at) = o;
    }
}

}  // namespace

extern "C" {

int fsx_clahe_workspace_bytes(int n_img, int tiles_x, int tiles_y, size_t* bytes) {
    FSX_REQUIRE(n_img >= 0 && tiles_x >= 1 && tiles_y >= 1 && tiles_x <= 256 && tiles_y <= 256 && bytes,
                "fsx_clahe_workspace_bytes: bad n_img=%d tiles=%dx%d", n_img, tiles_x, tiles_y);
    *bytes = (size_t)n_img * tiles_x * tiles_y * 256 * (1 + sizeof(unsigned));   // LUTs + histograms
    return FSX_OK;
}

int fsx_clahe(int n_img, const uint8_t* src_dev, int H, int W, double clip_limit, int tiles_x, int tiles_y,
              uint8_t* dst_dev, void* ws_dev, size_t ws_bytes, void* stream) {
    FSX_REQUIRE(n_img >= 0 && tiles_x >= 1 && tiles_y >= 1 && tiles_x <= 256 && tiles_y <= 256 && H > tiles_y &&
                    W > tiles_x && H <= 65535 && n_img <= 65535,
                "fsx_clahe: bad n_img=%d H=%d W=%d tiles=%dx%d", n_img, H, W, tiles_x, tiles_y);
    if (n_img == 0) return FSX_OK;
    FSX_REQUIRE(src_dev && dst_dev && ws_dev, "fsx_clahe: null pointer");
    const size_t n_bins = (size_t)n_img * tiles_x * tiles_y * 256;
    FSX_REQUIRE(ws_bytes >= n_bins * (1 + sizeof(unsigned)), "fsx_clahe: workspace %zu too small", ws_bytes);
    const bool divides = W % tiles_x == 0 && H % tiles_y == 0;
    const int We = divides ? W : W + tiles_x - W % tiles_x;
    const int He = divides ? H : H + tiles_y - H % tiles_y;
    const int tw = We / tiles_x, th = He / tiles_y;
    const int area = tw * th;
    int limit = 0;
    if (clip_limit > 0.0) {
        limit = (int)(clip_limit * area / 256);
        if (limit < 1) limit = 1;
    }
    const float scale = 255.0f / (float)area;
    const float inv_tw = 1.0f / (float)tw, inv_th = 1.0f / (float)th;
    uint8_t* lut = (uint8_t*)ws_dev;
    unsigned* hist = (unsigned*)(lut + ((n_bins + 255) & ~(size_t)255));
    FSX_REQUIRE(ws_bytes >= ((n_bins + 255) & ~(size_t)255) + n_bins * sizeof(unsigned),
                "fsx_clahe: workspace %zu too small", ws_bytes);
    // row slabs per tile: about 4096 histogram blocks in all
    const long long tiles_all = (long long)tiles_x * tiles_y * n_img;
    int slabs = (int)((4096 + tiles_all - 1) / tiles_all);
    slabs = slabs < 1 ? 1 : (slabs > th ? th : slabs);
    const int slab_rows = (th + slabs - 1) / slabs;
    slabs = (th + slab_rows - 1) / slab_rows;
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_clahe_zero, dim3(512), dim3(CL_THREADS), 0, st, hist, (long long)n_bins);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_clahe_hist, dim3(tiles_x * tiles_y, n_img, slabs), dim3(CL_THREADS), 0, st, src_dev, H, W,
                       tiles_x, th, tw, slab_rows, hist);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_clahe_lut, dim3(tiles_x * tiles_y, n_img), dim3(CL_THREADS), 0, st, (const unsigned*)hist,
                       limit, scale, lut);
    FSX_LAUNCH_CHECK();
    const bool vec = W % 4 == 0 && (((uintptr_t)src_dev | (uintptr_t)dst_dev) & 3) == 0;
    if (vec && tiles_x <= CL_LDS_TX && th >= CL_ROWS)
        hipLaunchKernelGGL(k_clahe_apply_lds, dim3((W / 4 + CL_THREADS - 1) / CL_THREADS, (H + CL_ROWS - 1) / CL_ROWS,
                                                   n_img),
                           dim3(CL_THREADS), 0, st, src_dev, H, W, tiles_x, tiles_y, inv_tw, inv_th,
                           (const uint8_t*)lut, dst_dev);
    else if (vec)
        hipLaunchKernelGGL(k_clahe_apply<4>, dim3((W / 4 + CL_THREADS - 1) / CL_THREADS, H, n_img), dim3(CL_THREADS), 0,
                           st, src_dev, H, W, tiles_x, tiles_y, inv_tw, inv_th, (const uint8_t*)lut, dst_dev);
    else
        hipLaunchKernelGGL(k_clahe_apply<1>, dim3((W + CL_THREADS - 1) / CL_THREADS, H, n_img), dim3(CL_THREADS), 0,
                           st, src_dev, H, W, tiles_x, tiles_y, inv_tw, inv_th, (const uint8_t*)lut, dst_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
