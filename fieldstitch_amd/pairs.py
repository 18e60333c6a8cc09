"""Candidate image pairs from GPS positions on the GPU (replaces the pair selection of
app/planesite/pipeline.py:452-476).

For photo i: its ``neighbors`` nearest photos (Euclidean distance of the GPS x, y) closer than
``2 * half_long[i] * 0.95`` (no limit when no photo has a relative altitude), plus the
temporal neighbours i+1 and i+2 closer than ``2 * half_long[i]``; pairs (min, max), unique,
sorted.  ``half_long[i] = agl[i] * max(W, H) / 2 / f_med`` is the ground half-extent of the
photo's long side (pipeline.py:459-460).  Exact ties in distance at the k-th neighbour go to
the lower index.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


def half_extents(agl, W, H, f_med):
    """pipeline.py:459-460: (half_short, half_long) ground half-extents per photo (host float64;
    agl, W, H per photo or scalars)."""
    agl = np.asarray(agl, np.float64)
    half_short = agl * (np.minimum(W, H) / 2.0) / f_med
    half_long = agl * (np.maximum(W, H) / 2.0) / f_med
    return half_short, half_long


def select_pairs(xy, half_long, neighbors, has_rel_any=True, stream=None):
    """Sorted list of (i, j), i < j (pipeline.py:453-476).  xy: (n, 2) GPS positions (float64),
    half_long: (n,) float64; the k-nearest search runs on the GPU."""
    import torch
    xy = np.ascontiguousarray(np.asarray(xy, np.float64)[:, :2])
    n = len(xy)
    if n < 2:
        return []
    k = min(n - 1, int(neighbors))
    hl = np.array(np.broadcast_to(np.asarray(half_long, np.float64), (n,)))
    d_xy = torch.from_numpy(xy).cuda()
    d_hl = torch.from_numpy(hl).cuda()
    out = torch.empty((n, k + 2, 2), dtype=torch.int32, device=d_xy.device)
    _lib.call("fsx_pair_select", n, C.c_void_p(d_xy.data_ptr()), C.c_void_p(d_hl.data_ptr()), k,
              1 if has_rel_any else 0, C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle(stream)))
    o = out.cpu().numpy().reshape(-1, 2)
    o = o[o[:, 0] >= 0]
    if len(o) == 0:
        return []
    o = np.unique(o, axis=0)      # sorted lexicographically, as sorted(set(...))
    return [(int(a), int(b)) for a, b in o]
