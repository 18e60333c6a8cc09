"""CPU oracle of contrast-limited adaptive histogram equalisation (test infrastructure only).

The reference conditions every photo with ``cv2.createCLAHE(clipLimit=2.0,
tileGridSize=(8, 8)).apply(gray)`` (app/planesite/features.py:60-71, :197, :225).  OpenCV
(cv2, 4.x, imgproc clahe.cpp, 8-bit path) is not installed here, so this module restates its
published algorithm step by step in numpy, float32 where OpenCV computes in float:

1. if W or H is not a multiple of the tile count, the image is extended at the right and
   bottom by (tiles - size % tiles) pixels each -- a whole tile-width / -height when that
   side already divides -- with BORDER_REFLECT_101; tile size = extended size / tiles;
2. per tile a 256-bin histogram; with clip > 0 each bin is capped at
   limit = max(int(clip * tile_area / 256), 1), the excess summed, clipped // 256 added to every
   bin and the remaining r = clipped % 256 added one each to bins 0, s, 2s, ... (r of them,
   s = max(256 // r, 1));
3. LUT[v] = round-half-even(float32(cumsum[v]) * float32(255 / tile_area)), clamped to 0..255;
4. pixel (x, y) of value v: txf = float32(x) * (1 / tile_w) - 0.5, tx1 = floor(txf),
   xa = txf - tx1, tiles clamped to the grid, the same along y, and
   out = round-half-even((L[ty1,tx1][v] * (1 - xa) + L[ty1,tx2][v] * xa) * (1 - ya) +
                         (L[ty2,tx1][v] * (1 - xa) + L[ty2,tx2][v] * xa) * ya), each float32
   operation rounded on its own.

Parity with cv2 itself is unpinned (no cv2, no reference-held vectors).  The restatement is
pinned in tests/test_oracle_clahe.py by what the mathematics fixes: one tile without clipping
is global histogram equalisation, the clip / redistribution of a hand-built histogram,
LUT[255] = 255 and monotone LUTs, identical tiles giving the tile LUT everywhere, and
BORDER_REFLECT_101 = numpy's "reflect".
"""
from __future__ import annotations

import numpy as np


def extend(img, tiles_x=8, tiles_y=8):
    """Step 1: the image the tiles are cut from."""
    H, W = img.shape
    if W % tiles_x == 0 and H % tiles_y == 0:
        return img
    return np.pad(img, ((0, tiles_y - H % tiles_y), (0, tiles_x - W % tiles_x)), mode="reflect")


def clip_histogram(hist, limit):
    """Step 2 on one histogram (int64, 256 bins)."""
    hist = hist.astype(np.int64).copy()
    clipped = int(np.maximum(hist - limit, 0).sum())
    hist = np.minimum(hist, limit)
    batch = clipped // 256
    residual = clipped - batch * 256
    hist += batch
    if residual:
        step = max(256 // residual, 1)
        hist[np.arange(0, 256, step)[:residual]] += 1
    return hist


def luts(img, clip=2.0, tiles_x=8, tiles_y=8):
    """Steps 1-3: (tiles_y, tiles_x, 256) uint8 LUTs and the tile size (th, tw)."""
    ext = extend(img, tiles_x, tiles_y)
    th, tw = ext.shape[0] // tiles_y, ext.shape[1] // tiles_x
    area = th * tw
    limit = max(int(clip * area / 256), 1) if clip > 0 else 0
    scale = np.float32(255) / np.float32(area)
    out = np.zeros((tiles_y, tiles_x, 256), np.uint8)
    for ty in range(tiles_y):
        for tx in range(tiles_x):
            hist = np.bincount(ext[ty * th:(ty + 1) * th, tx * tw:(tx + 1) * tw].ravel(), minlength=256)
            if clip > 0:
                hist = clip_histogram(hist, limit)
            v = np.cumsum(hist).astype(np.float32) * scale
            out[ty, tx] = np.clip(np.rint(v), 0, 255).astype(np.uint8)
    return out, (th, tw)


def _axis(n, t, tiles):
    inv = np.float32(1) / np.float32(t)
    f = np.arange(n).astype(np.float32) * inv - np.float32(0.5)
    i1 = np.floor(f).astype(np.int64)
    a = (f - i1.astype(np.float32)).astype(np.float32)
    return np.maximum(i1, 0), np.minimum(i1 + 1, tiles - 1), a, (np.float32(1) - a).astype(np.float32)


def apply(img, clip=2.0, tiles_x=8, tiles_y=8):
    """cv2.createCLAHE(clip, (tiles_x, tiles_y)).apply(img) for uint8 img, restated."""
    img = np.ascontiguousarray(img, np.uint8)
    L, (th, tw) = luts(img, clip, tiles_x, tiles_y)
    H, W = img.shape
    x1, x2, xa, xa1 = _axis(W, tw, tiles_x)
    y1, y2, ya, ya1 = _axis(H, th, tiles_y)
    v = img.astype(np.int64)
    f = lambda ty, tx: L[ty[:, None], tx[None, :], v].astype(np.float32)
    top = f(y1, x1) * xa1[None, :] + f(y1, x2) * xa[None, :]
    bot = f(y2, x1) * xa1[None, :] + f(y2, x2) * xa[None, :]
    res = top * ya1[:, None] + bot * ya[:, None]
    return np.clip(np.rint(res), 0, 255).astype(np.uint8)
