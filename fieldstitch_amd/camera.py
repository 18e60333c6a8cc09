"""Brown-Conrady undistortion of keypoint positions on the GPU (replaces app/planesite/camera.py
``undistort_pixels`` as used by matching._get, matching.py:48-64, and the rescale of
matching.load_features, matching.py:29-35).

``k`` = (f, cx, cy, k1, k2, p1, p2, k3) with fx = fy = f (OpenCV convention).  The device kernel
evaluates the same float64 expressions in the same order as the reference's numpy code
(25 fixed-point iterations, |rad| clamped at 0.2, early exit when every coefficient is zero),
so results are bit-identical.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

ITERS = 25


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("fieldstitch_amd.camera needs a ROCm GPU")
    return torch


def undistort_batch(items, iters=ITERS, want_xy=True, stream=None):
    """Batched rescale + undistortion on the device.

    items: list of (xy float32 (n,2) device tensor in reduced-image pixels, (sx, sy) scale to
    full resolution (W/rw, H/rh), k 8-tuple).  Returns a list of (xy float64 (n,2) full-resolution
    pixels or None, und float64 (n,2)) device tensors.  Asynchronous on ``stream``."""
    torch = _torch()
    n_img = len(items)
    if n_img == 0:
        return []
    dev = None
    xs, outs_xy, outs_und, scale, cam = [], [], [], [], []
    for xy, (sx, sy), k in items:
        if not (isinstance(xy, torch.Tensor) and xy.is_cuda and xy.dtype == torch.float32):
            raise TypeError("xy must be a float32 CUDA tensor")
        xy = xy.reshape(-1, 2).contiguous()
        dev = xy.device
        xs.append(xy)
        outs_und.append(torch.empty((xy.shape[0], 2), dtype=torch.float64, device=dev))
        outs_xy.append(torch.empty((xy.shape[0], 2), dtype=torch.float64, device=dev) if want_xy else None)
        scale += [float(sx), float(sy)]
        k = [float(v) for v in k]
        if len(k) != 8:
            raise ValueError("k must be (f, cx, cy, k1, k2, p1, p2, k3)")
        cam += k
    tb = C.c_size_t()
    _lib.call("fsx_undistort_table_bytes", n_img, C.byref(tb))
    table = torch.empty(tb.value, dtype=torch.uint8, device=dev)
    P = C.c_void_p
    nk = (C.c_int * n_img)(*[x.shape[0] for x in xs])
    c_xy = (P * n_img)(*[x.data_ptr() for x in xs])
    c_sc = (C.c_double * (2 * n_img))(*scale)
    c_cam = (C.c_double * (8 * n_img))(*cam)
    c_xo = (P * n_img)(*[o.data_ptr() for o in outs_xy]) if want_xy else None
    c_uo = (P * n_img)(*[o.data_ptr() for o in outs_und])
    _lib.call("fsx_undistort_batch", n_img, nk, c_xy, c_sc, c_cam, int(iters), c_xo, c_uo, P(table.data_ptr()),
              tb.value, P(_lib.stream_handle(stream)))
    if stream is not None:       # the table was allocated on the current stream
        table.record_stream(stream)
    return list(zip(outs_xy, outs_und))


def undistort_pixels(uv, k, iters=ITERS):
    """camera.py:28-30: distorted pixels (n,2) -> ideal pinhole pixels with the same f, cx, cy
    (float64 host array, computed on the GPU).  The float64 input is used as is (scale 1)."""
    torch = _torch()
    uv = np.asarray(uv, np.float64).reshape(-1, 2)
    n = len(uv)
    if n == 0:
        return np.zeros((0, 2), np.float64)
    d_uv = torch.from_numpy(np.ascontiguousarray(uv)).cuda()
    out = torch.empty((n, 2), dtype=torch.float64, device=d_uv.device)
    c_cam = (C.c_double * 8)(*[float(v) for v in k])
    _lib.call("fsx_undistort_f64", n, C.c_void_p(d_uv.data_ptr()), c_cam, int(iters), C.c_void_p(out.data_ptr()),
              C.c_void_p(_lib.stream_handle()))
    return out.cpu().numpy()
