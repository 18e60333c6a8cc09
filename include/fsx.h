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

/* Launch-shape tunables, process-wide (no reference counterpart; set by
 * fieldstitch_amd.autotune): key 0 = orientation blocks per image (16..65536, default 4096),
 * 1 = refinement blocks per image (16..65536, default 256), 2 = float matcher targets per LDS
 * tile (64 or 128, default 64), 3 = float matcher blocks of 32 queries per wave (1 or 2,
 * default 1), 4 = pyramid blur output rows per tile, the row strip (8, 16, 24 or 32, default
 * 16), 5 = guided matcher contraction workgroups (0 = one per chunk of 32 queries, the
 * default, or 256..65536 persistent ones striding over every pair's chunks).
 * Every value gives identical results.  Not synchronised with running work: set between
 * batches (a captured graph keeps the shapes it was captured with). */
int fsx_tune_set(int key, int value);
int fsx_tune_get(int key, int *value);

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

/* Bytes of the per-image device workspace (the reference's _Atlas, gsift.py:227-271, as the
 * six Gaussian layers of every octave: its DoG layers are their differences, formed by the
 * kernels that read them; plus keypoint buffers).  cand_cap: extrema candidate capacity;
 * kp_cap: oriented keypoint capacity, a power of two >= 1024. */
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
 * u8_dev (may be NULL): RootSIFT uint8 encoding of the same rows (features.py:80-82 _save).
 * A selection outside [0, kp_stride), or a keypoint whose octave / layer the workspace does
 * not hold, yields a zero row (the device cannot report it; gsift.describe checks both). */
int fsx_sift_describe(const void *ws_dev, const float *kp_dev, int kp_stride, const long long *sel_dev, int m,
                      float *desc_dev, uint8_t *u8_dev, void *stream);

/* Bytes of the device job table fsx_sift_features_batch needs for n_img images. */
int fsx_sift_table_bytes(int n_img, size_t *bytes);

/* The GPU feature worker for a batch of photos in one call (features.py:115-129 sift_gpu for
 * every photo, as gpu_worker_main's loop does one by one, features.py:343-378): detect
 * int(nfeat*2.2)=max_kp keypoints, grid-spread to nfeat (features.py:38-57 _spread), sort by
 * response, describe, RootSIFT-encode (features.py:80-82).  Per image i: imgs[i] (device
 * uint8, H[i] x W[i], pitch ld[i]), its workspace ws[i], detect arrays kp_out[i][7][max_kp],
 * n_out_dev[i]; descriptors in selection order: desc_out[i] (nfeat x 128 float32, array may be
 * NULL) and u8_out[i] (nfeat x 128 uint8, array may be NULL).  The selected count is
 * min(n_out, nfeat); the selection (int64, final order) sits in ws[i] at
 * fsx_sift_selection_offset.  Images of equal shape share their pyramid launches.  table_dev:
 * fsx_sift_table_bytes(n_img) bytes of device memory.  stage_events (may be NULL): 8 caller-created
 * hipEvent_t recorded on `stream` around the stages pyramid / extrema / refine / orientation /
 * select / spread / describe (asynchronous; the features.py timings of the reference). */
int fsx_sift_features_batch(int n_img, const uint8_t *const *imgs, const int *H, const int *W, const int *ld,
                            double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                            void *const *ws, const size_t *ws_bytes, float *const *kp_out, int *n_out_dev,
                            float *const *desc_out, uint8_t *const *u8_out, void *table_dev, size_t table_bytes,
                            void *const *stage_events, void *stream);

/* fsx_sift_features_batch in two halves for repeated runs on the same buffers (and for graph
 * capture): _prepare uploads the job records (asynchronous on `stream`), _launch only enqueues
 * the kernels (no host->device copy, no allocation, no memset).  Same arguments as the batch
 * call; the image / workspace / output pointers are baked in by _prepare.  No reference
 * counterpart. */
int fsx_sift_features_prepare(int n_img, const uint8_t *const *imgs, const int *H, const int *W, const int *ld,
                              double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                              void *const *ws, const size_t *ws_bytes, float *const *kp_out, int *n_out_dev,
                              float *const *desc_out, uint8_t *const *u8_out, void *table_dev, size_t table_bytes,
                              void *stream);
int fsx_sift_features_launch(int n_img, const uint8_t *const *imgs, const int *H, const int *W, const int *ld,
                             double contrast, double edge, int max_kp, int nfeat, int cand_cap, int kp_cap,
                             void *const *ws, const size_t *ws_bytes, float *const *kp_out, int *n_out_dev,
                             float *const *desc_out, uint8_t *const *u8_out, void *table_dev, size_t table_bytes,
                             void *const *stage_events, void *stream);

/* Byte offset of the spread selection inside a workspace (no reference counterpart). */
int fsx_sift_selection_offset(int H, int W, int cand_cap, int kp_cap, size_t *off);

/* ---------------------------------------------------------------------------------------
 * Matcher (replaces the matcher core of app/planesite/matching.py)
 * ------------------------------------------------------------------------------------- */

/* Workspace bytes of a batch of n_pair pairs (pa[p] -> pb[p]) over n_img keypoint sets
 * (n_kp[i] keypoints of an image W[i] x H[i]); out_capacity: total output slots
 * (sum of min(n_a, n_b)). */
int fsx_match_workspace_bytes(int n_img, const int *n_kp, const int *W, const int *H, int n_pair, const int *pa,
                              const int *pb, double radius, size_t *bytes, long long *out_capacity);

/* Batched matching (the match_pair pool of pipeline.py:535-548 over matching.py's core):
 * for each pair p, queries = keypoints of image pa[p], targets = keypoints of pb[p].
 *   dense[p] = 1: brute-force ratio matching of all queries (_ratio_match_bf,
 *     matching.py:86-100);
 *   dense[p] = 0: guided matching (_guided, matching.py:113-152): prediction
 *     Hm[9p..9p+8] applied to the query positions with cv2.perspectiveTransform semantics,
 *     queries predicted strictly inside (-radius, W+radius) x (-radius, H+radius), candidates
 *     strictly within radius (squared float64 distance, cKDTree's test), the kc spatially nearest
 *     (kc = 0: all; ties by index).
 * Then top-2, ratio test sqrt(b1) < ratio*sqrt(b2) in float64 (one candidate: always
 * accepted) and one query per target (matching.py:103-110, ties to the lower query), output
 * sorted by target.  desc_dev: all keypoint sets' uint8 descriptors concatenated (n x 128);
 * xy_dev: their float64 positions (n x 2).  form 0 = plain FP32 gather + VALU dot products,
 * form 1 = spatially blocked BF16 MFMA with fused window mask and top-2.  Both return the
 * FP64-oracle set (uint8 distances are exact integers).  out_offsets (host, n_pair): output
 * slot of each pair.  stage_events (may be NULL): 4 caller hipEvent_t recorded around
 * setup (grid bucketing) / contraction / compaction.  Asynchronous; read with
 * fsx_match_results.  Calls on one workspace must not overlap. */
int fsx_match_batch(int n_img, const int *n_kp, const int *W, const int *H, const uint8_t *desc_dev,
                    const double *xy_dev, int n_pair, const int *pa, const int *pb, const int *dense,
                    const double *Hm, double radius, int kc, double ratio, int form, void *ws_dev, size_t ws_bytes,
                    long long *out_offsets, void *stage_events, void *stream);

/* fsx_match_batch split in two for repeated runs of the same batch (and for graph capture):
 * fsx_match_prepare plans the batch and uploads its tables into the workspace (asynchronous
 * on `stream`; the device pointers and parameters are baked in), filling launch_info[64]
 * (host memory, opaque);
 * fsx_match_launch then only enqueues the kernels (no host->device copies, no allocation),
 * so it can be captured into a hipGraph and replayed.  No reference counterpart (the
 * reference re-plans per pair in a process pool). */
int fsx_match_prepare(int n_img, const int *n_kp, const int *W, const int *H, const uint8_t *desc_dev,
                      const double *xy_dev, int n_pair, const int *pa, const int *pb, const int *dense,
                      const double *Hm, double radius, int kc, double ratio, void *ws_dev, size_t ws_bytes,
                      long long *out_offsets, long long *launch_info, void *stream);
int fsx_match_launch(void *ws_dev, const long long *launch_info, int form, void *stage_events, void *stream);

/* Copies the last batch's results to the host (synchronises): counts[n_pair], q and t at the
 * pair offsets (int32, as match_pair returns them, matching.py:249), flags[4] (flags[2] != 0:
 * an internal table overflowed). */
int fsx_match_results(int n_img, const int *n_kp, const int *W, const int *H, int n_pair, const int *pa,
                      const int *pb, double radius, const void *ws_dev, int *counts, int *q, int *t, int *flags,
                      void *stream);

/* The last batch's (q, t) lists packed on the device (asynchronous; no counterpart in the
 * reference, whose per-pair lists are host arrays from the start, matching.py:249): pair p's
 * counts[p] entries (as fsx_match_results reports them) to rows poff_dev[p] .. of q_dev / t_dev
 * (int32, cap rows each; a list that would pass cap is not written); poff_dev: n_pair int64,
 * normally the exclusive prefix sums of the counts.  launch_info: as filled by
 * fsx_match_prepare for ws_dev. */
int fsx_match_pack(const void *ws_dev, const long long *launch_info, const long long *poff_dev, long long cap,
                   int *q_dev, int *t_dev, void *stream);

/* fsx_match_batch for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  Same
 * match set and arguments as fsx_match_batch, but the call allocates nothing (no device,
 * pinned or heap memory, no hipMalloc / hipFree): the workspace is a 32 MiB array in the
 * library's device data (one copy per GPU, held until the process ends), and the results go
 * straight to the caller's device buffers: q_dev / t_dev (int32, sum over pairs of
 * min(n_a, n_b) entries, pair p's list at out_offsets[p]), counts_dev (int32, n_pair) and
 * flags_dev (int32, 4; flags_dev[2] != 0: an internal table overflowed).  Limits: n_img <= 64,
 * n_pair <= 256, and the batch's workspace (fsx_match_workspace_bytes minus its output
 * region) within 32 MiB; beyond them the call fails with FSX_EINVAL and a message.
 * Asynchronous on `stream` (no host wait).  Every call on one GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams or threads) race; serialise them. */
int fsx_match_batch_noalloc(int n_img, const int *n_kp, const int *W, const int *H, const uint8_t *desc_dev,
                            const double *xy_dev, int n_pair, const int *pa, const int *pb, const int *dense,
                            const double *Hm, double radius, int kc, double ratio, int form, int *q_dev,
                            int *t_dev, int *counts_dev, int *flags_dev, long long *out_offsets, void *stream);

/* Per-query state of the last batch (synchronises): accepted best target (-1 = none), its
 * distance b1 and its grid cell (-1 = predicted outside B); arrays of sum(n_a) entries, may be
 * NULL.  For tests and diagnostics (no reference counterpart). */
int fsx_match_query_state(int n_img, const int *n_kp, const int *W, const int *H, int n_pair, const int *pa,
                          const int *pb, double radius, const void *ws_dev, int *best, double *b1, int *qcell,
                          void *stream);

/* Byte offsets inside a matcher workspace of its per-query arrays (caller's query order,
 * queries of pair p at the prefix of n_a): offs[0] best candidate before the ratio test
 * (int32, -1 = none), [1] its distance and the second smallest (float32 pairs), [2] the
 * accepted target (int32, -1 = none), [3] the prediction's grid cell (int32).  The learned
 * match filter's inputs (no reference counterpart). */
int fsx_match_candidate_offsets(int n_img, const int *n_kp, const int *W, const int *H, int n_pair, const int *pa,
                                const int *pb, double radius, long long *offs);

/* Writes the 64 x 16 accumulator of one 32x32x16 BF16 MFMA on fixed small integer operands, to
 * check the operand/result lane layout the matcher assumes (no reference counterpart). */
int fsx_mfma_layout_probe(float *out_dev, void *stream);

/* Writes the 64 x 16 int32 accumulator of one 32x32x32 INT8 MFMA on fixed small operands, to
 * check the lane layout the matcher's MFMA form assumes (no reference counterpart). */
int fsx_mfma_i8_layout_probe(int *out_dev, void *stream);

/* Largest difference in float32 ulps between the device atan2f and the float32 rounding of
 * the float64 atan2 over n seeded gradient-like argument pairs and the zero / sign / equal
 * magnitude cases; *max_ulp_dev must be zero on entry.  Bounds the orientation kernel's fast
 * bin path (no reference counterpart). */
int fsx_atan2_probe(long long n, unsigned long long seed, unsigned *max_ulp_dev, void *stream);

/* Largest circular difference, in units of 1e-9 degree, between the descriptor's gradient
 * angle (a fitted odd polynomial for atan, sift.hip desc_angle_deg) and the float64 atan2 in
 * degrees on [0, 360), over n seeded gradient-like argument pairs: out[0]; out[1] = the same
 * for the atan2f form.  out[0..1] must be zero on entry.  Bounds the descriptor's angle error
 * (no reference counterpart). */
int fsx_desc_angle_probe(long long n, unsigned long long seed, unsigned *out, void *stream);

/* ---------------------------------------------------------------------------------------
 * Float-descriptor brute-force matcher (matching.py:86-110 on float descriptors)
 * ------------------------------------------------------------------------------------- */

/* Workspace bytes for n_pair brute-force pairs (pa[p] -> pb[p]) over n_set float descriptor
 * sets of n_kp[i] descriptors; out_capacity: sum of min(n_a, n_b). */
int fsx_fmatch_workspace_bytes(int n_set, const int *n_kp, int n_pair, const int *pa, const int *pb, size_t *bytes,
                               long long *out_capacity);

/* _ratio_match_bf + _unique (matching.py:86-110) for every pair, on float32 descriptors desc_dev
 * (all sets concatenated, n x 128).  Screening on MFMA with FP32 accumulation in precision 1 =
 * BF16 or 2 = FP8 e4m3 with per-descriptor scales, the K (2 or 8) best screened targets per
 * query rescored in float64, a certificate from a bound on the screening error, and an exact
 * float64 pass for the queries it cannot decide (K = 0: the exact pass for every query): the result is the FP64-oracle match set
 * (distance sum_k (a_k - b_k)^2 in float64, best by index on ties, sqrt(b1) < ratio*sqrt(b2),
 * one query per target).  out_offsets (host, n_pair): output slot of each pair.
 * stage_events (may be NULL): 5 caller hipEvent_t.  Asynchronous. */
int fsx_fmatch_batch(int n_set, const int *n_kp, const float *desc_dev, int n_pair, const int *pa, const int *pb,
                     double ratio, int precision, int K, void *ws_dev, size_t ws_bytes, long long *out_offsets,
                     void *stage_events, void *stream);

/* Results of the last fsx_fmatch_batch (synchronises): counts[n_pair], q / t (int32) at the pair
 * offsets, stats[4] = queries decided by the exact pass, max and mean |screened - float64|
 * distance over the rescored candidates, number of rescored candidates. */
int fsx_fmatch_results(int n_set, const int *n_kp, int n_pair, const int *pa, const int *pb, const void *ws_dev,
                       int *counts, int *q, int *t, double *stats, void *stream);

/* Writes the 64 x 16 accumulator of one 32x32x64 FP8 (e4m3) MFMA with unit scales on fixed
 * small operands, to check the format and lane layout the float matcher assumes (no
 * reference counterpart). */
int fsx_fp8_layout_probe(float *out_dev, void *stream);

/* fsx_fp8_layout_probe's operands with E8M0 scales per lane (A row l % 32 scaled by
 * 2^((l%32) % 5 - 2), B column l % 32 by 2^((l%32) % 3 - 1)): checks how the float matcher's
 * power-of-two per-descriptor scales enter the MFMA (no reference counterpart). */
int fsx_fp8_scale_probe(float *out_dev, void *stream);

/* The float matcher's FP8 quantization converts with the hardware e4m3 conversion; this
 * compares it with a software round-to-nearest-even e4m3 conversion on every float in
 * [-448, 448] (a zero code of either sign for -0 counts as equal): out_dev[0] (uint64) =
 * mismatches, out_dev[1] = the smallest mismatching bit pattern (all ones if none).  No
 * reference counterpart. */
int fsx_fp8_cvt_probe(unsigned long long *out_dev, void *stream);

/* ---------------------------------------------------------------------------------------
 * Learned match filter: BF16 MLP kernels (no reference counterpart; trained on the inlier
 * labels of the plain pipeline, used to prune candidates before the exact ratio test)
 * ------------------------------------------------------------------------------------- */

/* C = op(A) op(B), bf16 operands (uint16 bit patterns, row-major, leading dimensions in
 * elements), FP32 accumulation on MFMA.  ta / tb: operand stored transposed (op(A)[m][k] =
 * A[k*lda + m]; op(B)[k][n] = B[n*ldb + k]).  epi 0: C16 = bf16(relu(acc + bias[n]));
 * 1: C32 = acc; 2: C16 = bf16(acc where mask[m][n] > 0, else 0).  K % 32 == 0. */
int fsx_mlp_gemm(int ta, int tb, int epi, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                 float *C32, void *C16, int ldc, const float *bias, const void *mask, void *stream);

/* Scratch bytes of fsx_mlp_gemm_splitk for an M x N product in `splits` K chunks. */
int fsx_mlp_splitk_scratch(int M, int N, int splits, size_t *bytes);

/* C32 (M x N, leading dimension ldc) = op(A) op(B) as fsx_mlp_gemm's epilogue 1, with K split
 * into `splits` chunks computed by separate blocks into scratch and summed in chunk order
 * (deterministic): the weight gradients of a batch (small M x N, long K). */
int fsx_mlp_gemm_splitk(int ta, int tb, const void *A, int lda, const void *B, int ldb, int M, int N, int K,
                        float *C32, int ldc, float *scratch, int splits, void *stream);

/* Output layer over nb rows of H2 (bf16, 256 wide): p = sigmoid(H2 w3 + b3) (p_out may be
 * NULL); with labels y: dZ2 = ((p - y) / nb) w3 masked by H2 > 0 (bf16) and per-4-row partials
 * part[ceil(nb/4)][258] = (sum dz H2[:, j], sum dz, sum BCE loss). */
int fsx_mlp_head(const void *H2, const float *w3, const float *b3, const float *y, int nb, float *p_out, void *dZ2,
                 float *part, void *stream);

/* Column sums in a fixed order (bf16 != 0: bf16 input, else float32); scratch of
 * fsx_mlp_colsum_scratch bytes. */
int fsx_mlp_colsum_scratch(int nr, int ncol, size_t *bytes);
int fsx_mlp_colsum(const void *in, int bf16, int nr, int ld, int ncol, float *out, float *scratch, void *stream);

/* Adam step `step` (>= 1, bias-corrected) on n float32 weights; w16 (may be NULL): bf16 copy. */
int fsx_mlp_adam(float *w, float *m, float *v, const float *g, void *w16, int n, float lr, float beta1, float beta2,
                 float eps, int step, void *stream);

/* float32 -> bf16 (round to nearest even). */
int fsx_mlp_to_bf16(const float *in, void *out, int n, void *stream);

/* Fused inference of the match filter (no counterpart in the reference): p_out[r] =
 * sigmoid(relu(relu(X[r] W1 + b1) W2 + b2) . w3 + b3) for the n rows of X (bf16 [n][288],
 * 16-byte aligned), hidden activations rounded to bf16 as in fsx_mlp_gemm + fsx_mlp_head.  The
 * weights are first packed by fsx_mlp_infer_pack (W1 bf16 [288][256], W2 bf16 [256][256]) into
 * fsx_mlp_infer_pack_bytes bytes of 16-byte aligned device memory; pack again after they change. */
int fsx_mlp_infer_pack_bytes(size_t *bytes);
int fsx_mlp_infer_pack(const void *W1, const void *W2, void *packed, void *stream);
int fsx_mlp_infer(const void *X, int n, const void *packed, const float *b1, const float *b2, const float *w3,
                  const float *b3, float *p_out, void *stream);

/* Deterministic homography RANSAC (the verification step of match_pair, matching.py:226,
 * with reproducible sampling): pair p's matches are rows off_dev[p] .. off_dev[p+1] (int64
 * device prefix) of pa_dev / pb_dev (float64 (x, y) in image a / b); n_hyp hypotheses per pair
 * from splitmix64 counter draws, 4-point DLT in float64, inlier iff reprojection error < thr.
 * Writes inlier_dev (byte per match) and n_in_dev from the winning hypothesis, and H_dev (9
 * float64 per pair, zeros without a model): the winner refit by normalized least squares on its
 * inliers when it has at least 5 (as cv2.findHomography refines its RANSAC model), h33 = 1.
 * The homography and fundamental-matrix forms (here and in fsx_ransac_batch) take a
 * stream-ordered temporary for their hypotheses (hipMallocAsync / hipFreeAsync on `stream`, at
 * most 256 MB, or 73 n_hyp bytes when one pair needs more); the similarity takes none. */
int fsx_ransac_homography(int n_pair, const long long *off_dev, const double *pa_dev, const double *pb_dev, int n_hyp,
                          double thr, unsigned long long seed, uint8_t *inlier_dev, double *H_dev, int *n_in_dev,
                          void *stream);

/* Deterministic fundamental-matrix RANSAC (the role of cv2.findFundamentalMat in _verify,
 * matching.py:175-191): same layout and sampling as fsx_ransac_homography; normalized 8-point
 * algorithm (full-pivot elimination of the 8 x 9 system, no entry of F fixed, each solution
 * projected to rank 2 as cv2's models are), inlier iff the larger squared
 * point-to-epipolar-line distance < thr^2.  F_dev: 9 float64 per pair (row-major,
 * x_b^T F x_a = 0, rank 2). */
int fsx_ransac_fundamental(int n_pair, const long long *off_dev, const double *pa_dev, const double *pb_dev,
                           int n_hyp, double thr, unsigned long long seed, uint8_t *inlier_dev, double *F_dev,
                           int *n_in_dev, void *stream);

/* Deterministic similarity RANSAC (the role of cv2.estimateAffinePartial2D in _similarity,
 * matching.py:194-202): 2-point samples, M = [[a, -b, tx], [b, a, ty]] in M_dev[0..5] of each
 * pair's 9 float64, inlier iff |M x_a - x_b| < thr. */
int fsx_ransac_similarity(int n_pair, const long long *off_dev, const double *pa_dev, const double *pb_dev,
                          int n_hyp, double thr, unsigned long long seed, uint8_t *inlier_dev, double *M_dev,
                          int *n_in_dev, void *stream);

/* The learned filter's inputs for every query of the last matcher batch on ws (launch_info of
 * fsx_match_prepare): row g (bf16, 288 wide; zero when the query has no candidate) =
 * |a - b| / 255 (128), a * b / 255^2 (128), prior offset of the candidate / radius (2),
 * sqrt(b1) / 1024, sqrt(min(b2, 4e6)) / 1024, sqrt(b1 / b2) (0 when b2 is +inf),
 * log(size_b / size_a), wrapped angle difference / 180, radius / 100, then 24 zeros.
 * size_dev / angle_dev: float32 per keypoint in the batch's concatenated order. */
int fsx_match_filter_features(void *ws_dev, const long long *launch_info, const float *size_dev,
                              const float *angle_dev, void *X_dev, void *stream);

/* The matcher in stages, so that the learned match filter can prune between the contraction
 * and the one-query-per-target rule (no reference counterpart: the learned filter's pruning stage
 * ahead of matching.py:146-152's acceptance and the verification of :226-229).  On a workspace
 * prepared by fsx_match_prepare:
 *   fsx_match_contract: setup, kc-nearest prefilter and contraction (`form` as
 *     fsx_match_launch); every query's best candidate, distances and ratio-test outcome are
 *     stored, nothing is published yet;
 *   fsx_match_accepted: list_dev (2 int32 per entry, n_query entries of capacity) receives the
 *     queries that passed the ratio test as (global query, pair), in arbitrary order, and
 *     *count_dev their number;
 *   fsx_match_filter_rows: row r of X_dev (bf16, 288 wide, cap rows) = the filter input of list
 *     entry r (the layout of fsx_match_filter_features), for r < *count_dev;
 *   fsx_match_publish: entries with score_dev[r] >= tau (score_dev NULL: all) take part in the
 *     one-query-per-target rule, the rest are rejected; then the per-pair (q, t) lists are
 *     compacted for fsx_match_results.  cap >= *count_dev (n_query always is).
 * With score_dev NULL the result equals fsx_match_launch's.  Asynchronous. */
int fsx_match_contract(void *ws_dev, const long long *launch_info, int form, void *stream);
int fsx_match_accepted(void *ws_dev, const long long *launch_info, int *list_dev, int *count_dev, void *stream);
int fsx_match_filter_rows(void *ws_dev, const long long *launch_info, const float *size_dev, const float *angle_dev,
                          const int *list_dev, const int *count_dev, long long cap, void *X_dev, void *stream);
int fsx_match_publish(void *ws_dev, const long long *launch_info, const int *list_dev, const int *count_dev,
                      long long cap, const float *score_dev, float tau, void *stream);

/* RANSAC for pairs that each carry their own threshold and seed (device arrays of n_pair):
 * model 0 = homography (as fsx_ransac_homography), 1 = fundamental matrix, 2 = similarity.
 * Pair p draws hypothesis h from (seed_dev[p], 0, h), so its result does not depend on which
 * other pairs share the launch (match_pair per pair, matching.py:226,180,196). */
int fsx_ransac_batch(int model, int n_pair, const long long *off_dev, const double *pa_dev, const double *pb_dev,
                     int n_hyp, const double *thr_dev, const unsigned long long *seed_dev, uint8_t *inlier_dev,
                     double *M_dev, int *n_in_dev, void *stream);

/* fsx_ransac_batch with adaptive stopping, as cv2's estimators stop (cv2.findHomography's
 * confidence 0.995 / 2000 iterations at matching.py:226, findFundamentalMat's 0.999 / 10000 at
 * :180, estimateAffinePartial2D's 0.99 / 2000 at :196): at most max_hyp hypotheses per pair,
 * drawn in rounds of 256 until the inlier ratio w of the pair's best model so far gives
 * (1 - w^K)^(hypotheses drawn) <= 1 - confidence (sample size K = 4, 8, 2 for models 0, 1, 2),
 * evaluated by multiplications only (w = c / n, w^K by K - 1 products, the power by squarings
 * and products) so a restatement decides identically.  The winner is the best hypothesis of the
 * rounds drawn (most inliers, then the lowest index).  used_dev (optional, int per pair):
 * hypotheses drawn.  0 < confidence < 1.  Takes one stream-ordered temporary (flags and one
 * round of hypotheses). */
int fsx_ransac_batch_adaptive(int model, int n_pair, const long long *off_dev, const double *pa_dev,
                              const double *pb_dev, int max_hyp, double confidence, const double *thr_dev,
                              const unsigned long long *seed_dev, uint8_t *inlier_dev, double *M_dev, int *n_in_dev,
                              int *used_dev, void *stream);

/* Gather of matched positions (the host-side fancy indexing ua[q], ub[t] of matching.py:180,
 * 196, 226 for many pairs at once): out_dev[k] = xy_dev[idx_dev[k]] (two float64 each) for
 * k < n; an index outside [0, n_xy) gives NaN.  xy_dev and out_dev 16-byte aligned. */
int fsx_gather_xy(long long n, const int *idx_dev, const double *xy_dev, long long n_xy, double *out_dev,
                  void *stream);

/* The chain's lists of matches moved between its stages on the device (no counterpart in the
 * reference, whose lists stay host arrays, matching.py:180-233): list L's cnt_dev[L] rows from
 * src_off_dev[L] of q_src_dev / t_src_dev (int32 feature indices; with keep_dev, only rows
 * whose byte is non-zero, in order) to rows dst_off_dev[L] .. of q_dst_dev / t_dst_dev, and
 * optionally their positions xy_dev[base_a_dev[L] + q] / xy_dev[base_b_dev[L] + t] (float64
 * (x, y), NaN outside [0, n_xy)) to pa_dst_dev / pb_dst_dev; rows at or past dst_cap are
 * skipped; kept_dev[L] (optional) = rows taken.  Any output may be null; xy and the position
 * outputs 16-byte aligned.  Asynchronous. */
int fsx_gather_lists(int n_list, const long long *src_off_dev, const int *cnt_dev, const long long *dst_off_dev,
                     long long dst_cap, const long long *base_a_dev, const long long *base_b_dev,
                     const int *q_src_dev, const int *t_src_dev, const uint8_t *keep_dev, const double *xy_dev,
                     long long n_xy, int *q_dst_dev, int *t_dst_dev, double *pa_dst_dev, double *pb_dst_dev,
                     int *kept_dev, void *stream);

/* The least-squares similarity of _similarity (matching.py:194-202, cv2's refit on the RANSAC
 * inliers) for many pairs: pair p's rows off_dev[p] .. off_dev[p+1] of pa_dev / pb_dev, one
 * inlier byte per row; out_dev[4p..4p+3] = (angle rad, scale, tx, ty) of b ~ s R a + t, NaN
 * with fewer than 2 inliers or coincident points. */
int fsx_similarity_fit_batch(int n_pair, const long long *off_dev, const double *pa_dev, const double *pb_dev,
                             const uint8_t *inlier_dev, double *out_dev, void *stream);

/* ---------------------------------------------------------------------------------------
 * Keypoint geometry around the matcher
 * ------------------------------------------------------------------------------------- */

/* Bytes of the device parameter table fsx_undistort_batch needs for n_img keypoint sets. */
int fsx_undistort_table_bytes(int n_img, size_t *bytes);

/* Feature positions -> full-resolution undistorted pixels for n_img keypoint sets: the
 * rescale of matching.load_features (matching.py:29-35, (xy + 0.5) * scale - 0.5) and
 * camera.undistort_pixels (camera.py:5-30, called by matching._get, matching.py:55-57),
 * bit-exact float64.  Set i: n_kp[i] float32 (x, y) pairs at xy_dev[i] (reduced-image px),
 * scale[2i..2i+1] = (W/rw, H/rh), cam8[8i..8i+7] = (f, cx, cy, k1, k2, p1, p2, k3), `iters`
 * fixed-point iterations (the reference uses 25).  Writes xy_out_dev[i] (n x 2 float64
 * full-resolution px; the array may be NULL) and und_out_dev[i] (n x 2 float64).  table_dev:
 * fsx_undistort_table_bytes(n_img) bytes of device memory.  Asynchronous. */
int fsx_undistort_batch(int n_img, const int *n_kp, const float *const *xy_dev, const double *scale,
                        const double *cam8, int iters, double *const *xy_out_dev, double *const *und_out_dev,
                        void *table_dev, size_t table_bytes, void *stream);

/* camera.undistort_pixels (camera.py:28-30) of n float64 pixels uv_dev (n x 2) -> out_dev. */
int fsx_undistort_f64(int n, const double *uv_dev, const double *cam8, int iters, double *out_dev, void *stream);

/* matching._predict (matching.py:155-172): ground-plane prior of n undistorted pixels u_dev
 * (n x 2 float64) of camera i into camera j.  cams (host, 43 doubles, row-major):
 * inv(K_i)[9], R_i[9], C_i[3], zg_i, R_j[9], K_j[9], C_j[3].  Writes pred_dev (n x 2 float64)
 * and ok_dev (n bytes: t > 0 and z > 0).  Asynchronous. */
int fsx_ground_predict(int n, const double *u_dev, const double *cams, double *pred_dev, uint8_t *ok_dev,
                       void *stream);

/* match_pair's prior filter (matching.py:222-225, 238-241) for n_pair pairs at once: rows
 * off_dev[p] .. off_dev[p+1] (int64 prefix) of ua_dev / ub_dev (float64 undistorted pixels of
 * the matches in images i and j); cams_dev: 43 float64 per pair in fsx_ground_predict's layout;
 * tol_dev[p] (float64): keep_dev[r] = prediction valid and |pred - ub| < tol; tol < 0 keeps every
 * row (the pair has no attitude prior).  max_rows: the largest pair's row count.
 * Asynchronous. */
int fsx_prior_keep_batch(int n_pair, const long long *off_dev, const double *ua_dev, const double *ub_dev,
                         const double *cams_dev, const double *tol_dev, int max_rows, uint8_t *keep_dev,
                         void *stream);

/* Candidate pairs from GPS positions (pipeline.py:452-476).  xy_dev: n x 2 float64;
 * half_long_dev: n float64; k = min(n - 1, neighbors) <= 64; has_rel != 0 applies the
 * 2 * half_long * 0.95 limit.  out_dev (n x (k + 2) x 2 int32): per photo its k neighbour
 * pairs (min, max) then its 2 temporal pairs (i, i + off); (-1, -1) where none.  The caller
 * takes the sorted set of the valid entries.  Asynchronous. */
int fsx_pair_select(int n, const double *xy_dev, const double *half_long_dev, int k, int has_rel, int *out_dev,
                    void *stream);

/* Contrast-limited adaptive histogram equalisation of n_img uint8 photos of H x W (contiguous,
 * n x H x W), the conditioning of features.py:60-71 (cv2.createCLAHE(clip_limit, (tiles_x,
 * tiles_y)).apply, OpenCV's 8-bit algorithm).  Workspace: per-tile LUTs and histograms,
 * n_img * tiles_x * tiles_y * 256 * 5 bytes.  dst may equal src.  Requires H > tiles_y,
 * W > tiles_x, H <= 65535, n_img <= 65535.  Asynchronous. */
int fsx_clahe_workspace_bytes(int n_img, int tiles_x, int tiles_y, size_t *bytes);
int fsx_clahe(int n_img, const uint8_t *src_dev, int H, int W, double clip_limit, int tiles_x, int tiles_y,
              uint8_t *dst_dev, void *ws_dev, size_t ws_bytes, void *stream);

/* Feature tracks from verified match lists (recon.py:14-56, build_tracks(pairs, nfeat, n_img)).
 * Workspace bytes and the observation capacity min(n_nodes, 2 * n_edges) for n_nodes features
 * in all (n_nodes < 2^31) and n_edges matches. */
int fsx_tracks_workspace_bytes(long long n_nodes, long long n_edges, size_t *bytes, long long *obs_capacity);

/* offsets_dev: n_img + 1 int64, the exclusive cumulative feature counts (offsets[n_img] =
 * n_nodes).  Pair p matches image pair_ij_dev[2p] to pair_ij_dev[2p+1]; its matches are
 * q_dev[k], t_dev[k] (int32 feature indices) for pair_start_dev[p] <= k < pair_start_dev[p+1]
 * (int64, n_pair + 1 entries, last = n_edges).  Writes the observations sorted by track, then
 * node: tid_dev / img_dev / feat_dev (int64, out_capacity >= the workspace call's
 * obs_capacity), and result_dev (4 int64): [n_observations, n_tracks, index error, capacity
 * error].  An image or feature index out of range sets the index error; the outputs are then
 * undefined.  Asynchronous; deterministic. */
int fsx_build_tracks(int n_img, const long long *offsets_dev, long long n_nodes, int n_pair, const int *pair_ij_dev,
                     const long long *pair_start_dev, const int *q_dev, const int *t_dev, long long n_edges,
                     void *ws_dev, size_t ws_bytes, long long out_capacity, long long *tid_dev, long long *img_dev,
                     long long *feat_dev, long long *result_dev, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* FSX_H */
