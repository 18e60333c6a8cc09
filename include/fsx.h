/* libfsx — MI355X (gfx950) SIFT front end and descriptor matcher behind a plain C ABI.
 *
 * Conventions
 *   - Every function returns an int status: 0 = ok, 1 = invalid argument, 2 = HIP error,
 *     3 = capacity exceeded, 4 = out of memory.  fsx_last_error() gives the calling thread's
 *     message.  (The reference raises Python exceptions; the ctypes wrapper turns a non-zero
 *     status back into RuntimeError so its fallback logic keeps working.)
 *   - Pointers named *_dev / device buffers are device memory owned by the caller (torch
 *     tensors on the Python side); the library never allocates or frees them.
 *   - `stream` is a hipStream_t passed as void* (NULL = the default stream).  All work is
 *     enqueued on it; functions that return host data synchronise it.
 */
#ifndef FSX_H
#define FSX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Library version (no reference counterpart). */
int fsx_version(void);

/* Last error message of the calling thread; returns its length.  Replaces the reference's
 * exception text (features.py:153-154, matching.py:250-253). */
int fsx_last_error(char *buf, size_t len);

/* Device name / CU count / HBM bytes.  Replaces features.gpu_info (features.py:258-267). */
int fsx_device_info(int device, char *name, size_t len, int *n_cu, long long *hbm_bytes);

/* ---------------------------------------------------------------------------------------
 * SIFT (replaces app/planesite/gsift.py)
 * ------------------------------------------------------------------------------------- */

/* Octave shapes of the scale space of an H x W image (gsift.py:158-172 _pyramid's octave
 * count and stop rule).  rows/cols: arrays of >= 16 ints (may be NULL). */
int fsx_sift_plan(int H, int W, int *n_oct, int *rows, int *cols);

/* float32 Gaussian taps used for layer 0 (base blur) .. 5 (gsift.py:52-82 _gauss_kernel /
 * _Kernels).  out: >= 64 floats. */
int fsx_sift_taps(int layer, float *out, int *len);

/* Bytes of the per-image device workspace (the reference's _Atlas, gsift.py:227-271, plus
 * keypoint buffers).  cand_cap: extrema candidate capacity; kp_cap: oriented keypoint
 * capacity, a power of two >= 1024. */
int fsx_sift_workspace_bytes(int H, int W, int cand_cap, int kp_cap, size_t *bytes);

/* gsift.detect (gsift.py:435-477): pyramid, extrema, refinement, orientation and the
 * response-ordered top max_kp selection for one uint8 image (img_dev, row pitch ld).
 * Writes kp_out_dev[7][max_kp] = x, y, size, angle, response (float32), octave, layer
 * (int32 bit patterns) and *n_out_dev.  contrast/edge as gsift.detect.  stage_ms (host,
 * 5 floats, may be NULL) receives pyramid / extrema / refine / orientation / select times and
 * makes the call synchronous.  After the call the workspace holds the state describe()
 * needs.  Check fsx_sift_counters()[3] == 0 (no capacity overflow). */
int fsx_sift_detect(const uint8_t *img_dev, int H, int W, int ld, double contrast, double edge, int max_kp,
                    int cand_cap, int kp_cap, void *ws_dev, size_t ws_bytes, float *kp_out_dev, int *n_out_dev,
                    float *stage_ms, void *stream);

/* Pipeline counters of a workspace (synchronises the stream): [0] candidates, [1] refined,
 * [2] oriented keypoints, [3] overflow flags (bit 0 candidates, 1 refined, 2 oriented).
 * No reference counterpart (the reference's buffers grow on the host). */
int fsx_sift_counters(const void *ws_dev, int *out4, void *stream);

/* gsift.describe (gsift.py:480-525, _descr_flat :800-871): 128-d descriptors (L2 norm 512)
 * of keypoints sel_dev[0..m) of the arrays kp_dev[7][kp_stride] (layout of fsx_sift_detect's
 * output), sampled from the workspace of the detect call.  desc_dev: m x 128 float32.
 * u8_dev (may be NULL): RootSIFT uint8 encoding of the same rows (features.py:80-82 _save). */
int fsx_sift_describe(const void *ws_dev, const float *kp_dev, int kp_stride, const long long *sel_dev, int m,
                      float *desc_dev, uint8_t *u8_dev, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FSX_H */
