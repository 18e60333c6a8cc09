"""CPU restatement of match_pair's decision chain (test infrastructure only), matching.py:205-253
with the oracle's exact parts in the places of the GPU kernels:

* brute-force and guided ratio matching: the FP64 oracle (oracle.match_oracle), which for the
  pipeline's uint8 RootSIFT descriptors is the reference's float32 numpy matcher up to exact
  ratio ties (matching.py:86-152);
* the ground-plane prior: oracle.geom_oracle.ground_predict (matching.py:155-172);
* cv2's estimators: the deterministic RANSAC restated in oracle.ransac_oracle, each pair's
  stage drawing from its own seed (pair slot 0), as the batched GPU chain does, with adaptive
  stopping at the confidence and iteration cap of each cv2 call (findHomography's defaults
  0.995 / 2000, findFundamentalMat(..., 0.999, 10000), estimateAffinePartial2D's defaults
  0.99 / 2000);
* the similarity: that RANSAC's inliers fitted by numpy's least-squares solver.

Stage order, thresholds, counts and return tuples follow matching.py:205-253 line by line.
"""
from __future__ import annotations

import math

import numpy as np

from . import geom_oracle as go
from . import match_oracle as mo
from . import ransac_oracle as ro


H_CAP, H_CONF = 2000, 0.995      # cv2.findHomography defaults (matching.py:226)
F_CAP, F_CONF = 10000, 0.999     # cv2.findFundamentalMat(..., 0.999, 10000) (matching.py:180)
S_CAP, S_CONF = 2000, 0.99       # cv2.estimateAffinePartial2D defaults (matching.py:196)


def pair_seed(i, j, salt):
    return (0x5EED0000 ^ (int(i) * 0x9E3779B1) ^ (int(j) << 24) ^ salt) & 0xFFFFFFFFFFFFFFFF


def _prior_keep(ctx, i, j, ua, ub):
    """matching.py:221-225: prediction valid and within prior_tol * max(W_j, H_j)."""
    K, R, Cc, zg = ctx["K"], ctx["R"], ctx["C"], ctx["zg"]
    pred, ok = go.ground_predict(ua, np.linalg.inv(K[i]), R[i], Cc[i], zg[i], R[j], K[j], Cc[j])
    tol = ctx["prior_tol"] * max(ctx["W"][j], ctx["H"][j])
    return ok & (np.linalg.norm(pred - ub, axis=1) < tol)


def _verify(ctx, i, j, ua, ub, q, t):
    """matching.py:175-191 with the restated fundamental-matrix RANSAC."""
    if len(q) < 8:
        return None, None, 0
    f, _, cnt, _ = ro.ransac_pair_adaptive(0, ua[q], ub[t], int(ctx.get("ransac_iters", F_CAP)), F_CONF,
                                           float(ctx["ransac_thr"]), pair_seed(i, j, 0xF), "fundamental")
    if cnt == 0:
        return None, None, 0
    return q[f], t[f], int(f.sum())


def _similarity(ctx, i, j, ua, ub, q, t):
    """matching.py:194-202: similarity RANSAC, then the least-squares fit on its inliers."""
    if len(q) < 2:
        return None
    a, b = ua[q], ub[t]
    f, _, cnt, _ = ro.ransac_pair_adaptive(0, a, b, S_CAP, S_CONF, 0.02 * max(ctx["W"][i], ctx["H"][i]),
                                           pair_seed(i, j, 0x5), "similarity")
    if cnt < 2:
        return None
    a, b = a[f], b[f]
    A = np.zeros((2 * len(a), 4))
    A[0::2] = np.c_[a[:, 0], -a[:, 1], np.ones(len(a)), np.zeros(len(a))]
    A[1::2] = np.c_[a[:, 1], a[:, 0], np.zeros(len(a)), np.ones(len(a))]
    (sa, sb, tx, ty), *_ = np.linalg.lstsq(A, b.reshape(-1), rcond=None)
    return (math.atan2(sb, sa), math.hypot(sa, sb), float(tx), float(ty))


def match_pair(ctx, feats, i, j):
    """feats[k] = (desc uint8 (n, 128) in response order, und float64 (n, 2)).  Returns the
    reference's tuple for the pair (i, j)."""
    da, ua = feats[i]
    db, ub = feats[j]
    if len(ua) < 30 or len(ub) < 30:
        return i, j, None, None, None, 0
    use_prior = ctx["has_rot"][i] and ctx["has_rot"][j] and ctx.get("use_prior", True)
    result = None
    n1 = ctx.get("n_coarse", 1500)
    if ctx.get("guided", True) and len(ua) > n1 // 2:
        q1, t1, _ = mo.match_fp64_fast(da[:n1], db[:n1], 0.85)
        if use_prior and len(q1):
            m = _prior_keep(ctx, i, j, ua[q1], ub[t1])
            q1, t1 = q1[m], t1[m]
        if len(q1) >= 12:
            hm, Hm, cnt, _ = ro.ransac_pair_adaptive(0, ua[q1], ub[t1], H_CAP, H_CONF,
                                                     0.012 * max(ctx["W"][i], ctx["H"][i]), pair_seed(i, j, 0x1),
                                                     "homography")
            if cnt > 0 and hm.sum() >= 10:
                radius = ctx["guided_radius"] * max(ctx["W"][j], ctx["H"][j])
                kc = ctx.get("guided_k", 24)
                qs, cands = mo.guided_candidates(mo.predict(ua, Hm.reshape(3, 3)), ub, (ctx["W"][j], ctx["H"][j]),
                                                 radius, 0 if kc >= len(ub) else kc)
                q, t, _ = mo.match_fp64(da, db, 0.85, qs, cands)
                if len(q) >= ctx["min_inliers"]:
                    q, t, n_in = _verify(ctx, i, j, ua, ub, q, t)
                    if q is not None and n_in >= max(ctx["min_inliers"], 40):
                        result = (q, t, n_in)
    if result is None:
        q, t, _ = mo.match_fp64_fast(da, db, ctx.get("ratio", 0.8))
        if use_prior and len(q):
            m = _prior_keep(ctx, i, j, ua[q], ub[t])
            q, t = q[m], t[m]
        if len(q) < 20:
            return i, j, None, None, None, 0
        q, t, n_in = _verify(ctx, i, j, ua, ub, q, t)
        if q is None or n_in < ctx["min_inliers"]:
            return i, j, None, None, None, n_in
        result = (q, t, n_in)
    q, t, n_in = result
    return i, j, q.astype(np.int32), t.astype(np.int32), _similarity(ctx, i, j, ua, ub, q, t), n_in
