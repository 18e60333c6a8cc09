"""CPU oracle of the keypoint geometry around the matcher (test infrastructure only).

* ``rescale`` — matching.py:29-35 load_features: reduced-image positions to full resolution;
* ``undistort_pixels`` — camera.py:5-30 (Brown-Conrady, fx = fy, 25 fixed-point iterations,
  |rad| clamped at 0.2, early return for zero coefficients), numpy float64 in the reference's
  expression order;
* ``distort_pixels`` — the forward Brown-Conrady model (used only to pin the undistortion by
  a round trip);
* ``ground_predict`` — matching.py:155-172 _predict with explicit camera parameters;
* ``project`` — pinhole projection K R (G - C) (pins ground_predict by a closed form);
* ``select_pairs_kdtree`` — pipeline.py:452-476 with scipy's cKDTree, as the reference;
  ``select_pairs_brute`` — the same definition by brute force with ties to the lower index.
"""
from __future__ import annotations

import numpy as np


def rescale(xy32, W, H, rw, rh):
    xy = np.asarray(xy32).astype(np.float64)
    if len(xy) == 0:
        return xy.reshape(-1, 2)
    return (xy + 0.5) * np.array([W / float(rw), H / float(rh)]) - 0.5


def undistort_pixels(uv, k, iters=25):
    f, cx, cy, k1, k2, p1, p2, k3 = [np.float64(v) for v in k]
    uv = np.asarray(uv, np.float64).reshape(-1, 2)
    xd = (uv[:, 0] - cx) / f
    yd = (uv[:, 1] - cy) / f
    x, y = xd.copy(), yd.copy()
    if not (k1 == 0 and k2 == 0 and k3 == 0 and p1 == 0 and p2 == 0):
        for _ in range(iters):
            r2 = x * x + y * y
            rad = 1 + k1 * r2 + k2 * r2 * r2 + k3 * r2 * r2 * r2
            rad = np.where(np.abs(rad) < 0.2, 0.2, rad)
            tx = 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
            ty = p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
            x = (xd - tx) / rad
            y = (yd - ty) / rad
    return np.c_[x * f + cx, y * f + cy]


def distort_pixels(uv_ideal, k):
    """Forward model: ideal pinhole pixels -> distorted pixels."""
    f, cx, cy, k1, k2, p1, p2, k3 = [float(v) for v in k]
    uv = np.asarray(uv_ideal, np.float64).reshape(-1, 2)
    x = (uv[:, 0] - cx) / f
    y = (uv[:, 1] - cy) / f
    r2 = x * x + y * y
    rad = 1 + k1 * r2 + k2 * r2 ** 2 + k3 * r2 ** 3
    xd = x * rad + 2 * p1 * x * y + p2 * (r2 + 2 * x * x)
    yd = y * rad + p1 * (r2 + 2 * y * y) + 2 * p2 * x * y
    return np.c_[xd * f + cx, yd * f + cy]


def ground_predict(u, Kinv_i, R_i, C_i, zg_i, R_j, K_j, C_j):
    u = np.asarray(u, np.float64).reshape(-1, 2)
    rays = np.c_[u, np.ones(len(u))] @ np.asarray(Kinv_i).T
    d = rays @ np.asarray(R_i)
    dz = d[:, 2]
    dz = np.where(np.abs(dz) < 1e-6, -1e-6, dz)
    C_i = np.asarray(C_i, np.float64)
    t = (zg_i - C_i[2]) / dz
    G = C_i[None] + t[:, None] * d
    q = (G - np.asarray(C_j)[None]) @ np.asarray(R_j).T @ np.asarray(K_j).T
    z = q[:, 2]
    ok = (t > 0) & (z > 0)
    z = np.where(np.abs(z) < 1e-9, 1e-9, z)
    return q[:, :2] / z[:, None], ok


def project(G, K, R, Cc):
    q = (np.asarray(G, np.float64) - np.asarray(Cc)[None]) @ np.asarray(R).T @ np.asarray(K).T
    return q[:, :2] / q[:, 2:3], q[:, 2]


def _pairs_from(dists, nbrs, xy, half_long, has_rel_any):
    n = len(xy)
    pairs = set()
    for i in range(n):
        for d, j in zip(dists[i][1:], nbrs[i][1:]):
            if j >= n:
                continue
            lim = 2 * half_long[i] * 0.95 if has_rel_any else np.inf
            if d < lim:
                pairs.add((min(i, int(j)), max(i, int(j))))
    for i in range(n - 1):
        for off in (1, 2):
            j = i + off
            if j < n and np.hypot(*(xy[i] - xy[j])) < 2 * half_long[i]:
                pairs.add((i, j))
    return sorted(pairs)


def select_pairs_kdtree(xy, half_long, neighbors, has_rel_any=True):
    from scipy.spatial import cKDTree
    xy = np.asarray(xy, np.float64)[:, :2]
    n = len(xy)
    if n < 2:
        return []
    k = min(n - 1, neighbors)
    dists, nbrs = cKDTree(xy).query(xy, k=k + 1)
    return _pairs_from(dists, nbrs, xy, np.broadcast_to(half_long, (n,)), has_rel_any)


def select_pairs_brute(xy, half_long, neighbors, has_rel_any=True):
    xy = np.asarray(xy, np.float64)[:, :2]
    n = len(xy)
    if n < 2:
        return []
    k = min(n - 1, neighbors)
    dists = np.zeros((n, k + 1))
    nbrs = np.zeros((n, k + 1), np.int64)
    for i in range(n):
        d = np.sqrt((xy[:, 0] - xy[i, 0]) ** 2 + (xy[:, 1] - xy[i, 1]) ** 2)
        o = np.lexsort((np.arange(n), d))[:k + 1]
        dists[i], nbrs[i] = d[o], o
    return _pairs_from(dists, nbrs, xy, np.broadcast_to(half_long, (n,)), has_rel_any)
