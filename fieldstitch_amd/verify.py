"""Deterministic homography RANSAC on the GPU (the verification role of matching.py:226
cv2.findHomography(..., RANSAC) in match_pair, with reproducible sampling): labels the plain
pipeline's matches as geometric inliers."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def ransac_homography(pa_list, pb_list, n_hyp=512, thr=3.0, seed=0x5EED):
    """pa_list / pb_list: per pair (n_p, 2) float64 matched positions (numpy or device tensors).
    Returns (list of inlier bool arrays, H (n_pair, 3, 3) float64, counts (n_pair,))."""
    import torch
    n_pair = len(pa_list)
    if n_pair == 0:
        return [], np.zeros((0, 3, 3)), np.zeros(0, np.int64)
    t = lambda a: a.double() if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a, np.float64))  # noqa
    pa = [t(a).reshape(-1, 2) for a in pa_list]
    pb = [t(b).reshape(-1, 2) for b in pb_list]
    lens = [len(a) for a in pa]
    off = np.r_[0, np.cumsum(lens)].astype(np.int64)
    dev = "cuda"
    A = torch.cat(pa).to(dev).contiguous() if off[-1] else torch.zeros((1, 2), dtype=torch.float64, device=dev)
    B = torch.cat(pb).to(dev).contiguous() if off[-1] else torch.zeros((1, 2), dtype=torch.float64, device=dev)
    d_off = torch.from_numpy(off).to(dev)
    inl = torch.zeros(max(int(off[-1]), 1), dtype=torch.uint8, device=dev)
    H = torch.zeros((n_pair, 9), dtype=torch.float64, device=dev)
    cnt = torch.zeros(n_pair, dtype=torch.int32, device=dev)
    P = C.c_void_p
    _lib.call("fsx_ransac_homography", n_pair, P(d_off.data_ptr()), P(A.data_ptr()), P(B.data_ptr()), int(n_hyp),
              float(thr), C.c_ulonglong(seed), P(inl.data_ptr()), P(H.data_ptr()), P(cnt.data_ptr()),
              P(_lib.stream_handle()))
    flags = inl.cpu().numpy().astype(bool)
    return [flags[off[p]:off[p + 1]] for p in range(n_pair)], H.cpu().numpy().reshape(-1, 3, 3), cnt.cpu().numpy()
