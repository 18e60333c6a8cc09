"""Deterministic RANSAC on the GPU for the verification roles of match_pair: homography
(cv2.findHomography, matching.py:226), fundamental matrix (cv2.findFundamentalMat in _verify,
matching.py:175-191) and similarity (cv2.estimateAffinePartial2D in _similarity,
matching.py:194-202).  cv2's own sampling is not reproducible; these draw every sample from a
counter hash of (seed, pair, hypothesis), so a CPU restatement reproduces every count."""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

_ENTRY = {"homography": "fsx_ransac_homography", "fundamental": "fsx_ransac_fundamental",
          "similarity": "fsx_ransac_similarity"}


def ransac(model, pa_list, pb_list, n_hyp=512, thr=3.0, seed=0x5EED):
    """model: "homography" | "fundamental" | "similarity".  pa_list / pb_list: per pair (n_p, 2)
    float64 matched positions (numpy or device tensors).  Returns (list of inlier bool arrays,
    models (n_pair, 9) float64 (H / F row-major; similarity: a, -b, tx, b, a, ty, 0, 0, 1), all
    zero without a model, counts (n_pair,))."""
    import torch
    entry = _ENTRY[model]
    n_pair = len(pa_list)
    if n_pair == 0:
        return [], np.zeros((0, 9)), np.zeros(0, np.int64)
    t = lambda a: a.double() if isinstance(a, torch.Tensor) else torch.from_numpy(np.asarray(a, np.float64))  # noqa
    pa = [t(a).reshape(-1, 2) for a in pa_list]
    pb = [t(b).reshape(-1, 2) for b in pb_list]
    lens = [len(a) for a in pa]
    if any(len(a) != len(b) for a, b in zip(pa, pb)):
        raise ValueError("pa and pb must pair up row by row")
    off = np.r_[0, np.cumsum(lens)].astype(np.int64)
    dev = "cuda"
    A = torch.cat(pa).to(dev).contiguous() if off[-1] else torch.zeros((1, 2), dtype=torch.float64, device=dev)
    B = torch.cat(pb).to(dev).contiguous() if off[-1] else torch.zeros((1, 2), dtype=torch.float64, device=dev)
    d_off = torch.from_numpy(off).to(dev)
    inl = torch.zeros(max(int(off[-1]), 1), dtype=torch.uint8, device=dev)
    Mo = torch.zeros((n_pair, 9), dtype=torch.float64, device=dev)
    cnt = torch.zeros(n_pair, dtype=torch.int32, device=dev)
    P = C.c_void_p
    _lib.call(entry, n_pair, P(d_off.data_ptr()), P(A.data_ptr()), P(B.data_ptr()), int(n_hyp), float(thr),
              C.c_ulonglong(seed), P(inl.data_ptr()), P(Mo.data_ptr()), P(cnt.data_ptr()), P(_lib.stream_handle()))
    flags = inl.cpu().numpy().astype(bool)
    return [flags[off[p]:off[p + 1]] for p in range(n_pair)], Mo.cpu().numpy(), cnt.cpu().numpy()


def ransac_homography(pa_list, pb_list, n_hyp=512, thr=3.0, seed=0x5EED):
    """Homography RANSAC: (list of inlier flags, H (n_pair, 3, 3), counts)."""
    flags, Mo, cnt = ransac("homography", pa_list, pb_list, n_hyp, thr, seed)
    return flags, Mo.reshape(-1, 3, 3), cnt
