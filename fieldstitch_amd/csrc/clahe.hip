// Contrast-limited adaptive histogram equalisation of uint8 photos (the input conditioning of
// features.py:60-71: cv2.createCLAHE(clipLimit=2.0, tileGridSize=(8, 8)).apply(gray)).
//
// OpenCV's 8-bit algorithm, step for step (float32 where it computes in float):
//   * tiles are cut from the image extended at the right / bottom by (tiles - size % tiles)
//     pixels (a whole tile when that side divides) with BORDER_REFLECT_101, unless both sides
//     divide;
//   * k_clahe_lut, one block per (photo, tile): 256-bin histogram (one LDS copy per wave),
//     bins capped at `limit`, the excess spread as clipped / 256 per bin plus one each to
//     bins 0, s, 2s, ... (clipped % 256 of them, s = max(256 / that, 1)), inclusive scan,
//     LUT = round-half-even(float(cumsum) * (255 / tile area)) clamped to 0..255;
//   * k_clahe_apply, one thread per pixel: bilinear blend of the four surrounding tiles' LUT
//     values with weights from x * (1 / tile width) - 0.5 (and the same along y), each float
//     operation rounded on its own (the library is built without contraction).
#include "fsx_common.h"

namespace {

constexpr int CL_THREADS = 256;

__global__ void __launch_bounds__(CL_THREADS) k_clahe_lut(const uint8_t* src, int H, int W, int tiles_x, int th,
                                                          int tw, int limit, float scale, uint8_t* lut) {
    __shared__ unsigned s_hist[CL_THREADS / 64][256];
    __shared__ int s_wave[CL_THREADS / 64];
    const int tile = blockIdx.x, img = blockIdx.y;
    const int ty = tile / tiles_x, tx = tile - ty * tiles_x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < (CL_THREADS / 64) * 256; k += CL_THREADS) (&s_hist[0][0])[k] = 0;
    __syncthreads();
    const uint8_t* im = src + (size_t)img * H * W;
    for (int r = wave; r < th; r += CL_THREADS / 64) {
        int y = ty * th + r;
        if (y >= H) y = 2 * (H - 1) - y;                       // BORDER_REFLECT_101
        const uint8_t* row = im + (size_t)y * W;
        for (int c = lane; c < tw; c += 64) {
            int x = tx * tw + c;
            if (x >= W) x = 2 * (W - 1) - x;
            atomicAdd(&s_hist[wave][row[x]], 1u);
        }
    }
    __syncthreads();
    const int i = threadIdx.x;                                 // one bin per thread
    int h = 0;
    for (int w = 0; w < CL_THREADS / 64; ++w) h += (int)s_hist[w][i];
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
    lut[((size_t)img * gridDim.x + tile) * 256 + i] = (uint8_t)(v < 0.f ? 0.f : (v > 255.f ? 255.f : v));
}

__global__ void __launch_bounds__(CL_THREADS) k_clahe_apply(const uint8_t* src, int H, int W, int tiles_x, int tiles_y,
                                                            float inv_tw, float inv_th, const uint8_t* lut,
                                                            uint8_t* dst) {
    const int x = blockIdx.x * CL_THREADS + threadIdx.x;
    const int y = blockIdx.y, img = blockIdx.z;
    if (x >= W) return;
    const size_t at = ((size_t)img * H + y) * W + x;
    const int v = src[at];
    const float txf = (float)x * inv_tw - 0.5f;
    int tx1 = (int)floorf(txf);
    int tx2 = tx1 + 1;
    const float xa = txf - (float)tx1, xa1 = 1.0f - xa;
    tx1 = tx1 > 0 ? tx1 : 0;
    tx2 = tx2 < tiles_x - 1 ? tx2 : tiles_x - 1;
    const float tyf = (float)y * inv_th - 0.5f;
    int ty1 = (int)floorf(tyf);
    int ty2 = ty1 + 1;
    const float ya = tyf - (float)ty1, ya1 = 1.0f - ya;
    ty1 = ty1 > 0 ? ty1 : 0;
    ty2 = ty2 < tiles_y - 1 ? ty2 : tiles_y - 1;
    const uint8_t* L = lut + (size_t)img * tiles_x * tiles_y * 256;
    const uint8_t* p1 = L + (size_t)ty1 * tiles_x * 256;
    const uint8_t* p2 = L + (size_t)ty2 * tiles_x * 256;
    const float res = ((float)p1[tx1 * 256 + v] * xa1 + (float)p1[tx2 * 256 + v] * xa) * ya1 +
                      ((float)p2[tx1 * 256 + v] * xa1 + (float)p2[tx2 * 256 + v] * xa) * ya;
    const float r = rintf(res);
    dst[at] = (uint8_t)(r < 0.f ? 0.f : (r > 255.f ? 255.f : r));
}

}  // namespace

extern "C" {

int fsx_clahe_workspace_bytes(int n_img, int tiles_x, int tiles_y, size_t* bytes) {
    FSX_REQUIRE(n_img >= 0 && tiles_x >= 1 && tiles_y >= 1 && tiles_x <= 256 && tiles_y <= 256 && bytes,
                "fsx_clahe_workspace_bytes: bad n_img=%d tiles=%dx%d", n_img, tiles_x, tiles_y);
    *bytes = (size_t)n_img * tiles_x * tiles_y * 256;
    return FSX_OK;
}

int fsx_clahe(int n_img, const uint8_t* src_dev, int H, int W, double clip_limit, int tiles_x, int tiles_y,
              uint8_t* dst_dev, void* ws_dev, size_t ws_bytes, void* stream) {
    FSX_REQUIRE(n_img >= 0 && tiles_x >= 1 && tiles_y >= 1 && tiles_x <= 256 && tiles_y <= 256 && H > tiles_y &&
                    W > tiles_x && H <= 65535 && n_img <= 65535,
                "fsx_clahe: bad n_img=%d H=%d W=%d tiles=%dx%d", n_img, H, W, tiles_x, tiles_y);
    if (n_img == 0) return FSX_OK;
    FSX_REQUIRE(src_dev && dst_dev && ws_dev, "fsx_clahe: null pointer");
    FSX_REQUIRE(ws_bytes >= (size_t)n_img * tiles_x * tiles_y * 256, "fsx_clahe: workspace %zu too small", ws_bytes);
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
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_clahe_lut, dim3(tiles_x * tiles_y, n_img), dim3(CL_THREADS), 0, st, src_dev, H, W, tiles_x,
                       th, tw, limit, scale, (uint8_t*)ws_dev);
    FSX_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_clahe_apply, dim3((W + CL_THREADS - 1) / CL_THREADS, H, n_img), dim3(CL_THREADS), 0, st,
                       src_dev, H, W, tiles_x, tiles_y, inv_tw, inv_th, (const uint8_t*)ws_dev, dst_dev);
    FSX_LAUNCH_CHECK();
    return FSX_OK;
}

}  // extern "C"
