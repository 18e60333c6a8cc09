"""CPU oracle of the descriptor matcher (test infrastructure only).

Two layers, both plain numpy:

* a restatement of the reference's float32 matcher (app/planesite/matching.py): the
  brute-force ratio test ``_ratio_match_bf`` (:86-100), ``_unique`` (:103-110) and the
  prior-guided matcher ``_guided`` (:113-152), with cKDTree's k-nearest-within-radius replaced
  by a brute-force spatial search of the same definition;
* the FP64 oracle that defines the match set exactly (every GPU precision mode must return
  this list bit for bit):
    - candidates of query q: all of B (brute force), or the B keypoints strictly inside the
      prior radius around pred[q] (``radius`` in px), optionally only the kc spatially nearest
      (ties in spatial distance: lower index first);
    - d(q, t) = sum_k (a_k - b_k)^2 accumulated in float64 in k order;
    - best = argmin d (ties: lower t); b2 = the second smallest d (+inf if q has one
      candidate); accept iff sqrt(b1) < ratio * sqrt(b2) in float64;
    - one query per target: the smallest b1 wins, ties to the lower q; output sorted by t.

For uint8 RootSIFT descriptors (the pipeline's format) every distance is an exact integer
below 2^24, so the float32 restatement and the FP64 oracle differ only where
b1 / b2 == ratio^2 exactly; tests/test_oracle_match.py checks both against each other and
against scipy's cKDTree.
"""
from __future__ import annotations

import numpy as np


def _first_of_runs(s):
    """True at the first element of each run of equal values of s (empty for empty s)."""
    f = np.ones(len(s), bool)
    f[1:] = s[1:] != s[:-1]
    return f


def unique(q: np.ndarray, t: np.ndarray, dd: np.ndarray):
    """matching.py:103-110: keep one query per target (smallest dd, then earliest position);
    output sorted by target."""
    if len(q) == 0:
        return q, t
    order = np.lexsort((dd, t))
    t_s = t[order]
    first = _first_of_runs(t_s)
    keep = order[first]
    return q[keep], t[keep]


def _two_smallest(d2: np.ndarray):
    """Row-wise (best index, best value, second value); ties -> lower index (values are what
    the reference's argpartition pair yields: a tie at the minimum always fails the ratio test,
    so its index choice never reaches the output)."""
    n = d2.shape[1]
    best = np.argmin(d2, axis=1)
    r = np.arange(d2.shape[0])
    b1 = d2[r, best]
    if n < 2:
        return best, b1, np.full(len(r), np.inf, d2.dtype)
    tmp = d2.copy()
    tmp[r, best] = np.inf
    b2 = tmp.min(axis=1)
    return best, b1, b2


def ratio_match_bf_f32(da: np.ndarray, db: np.ndarray, ratio: float, na=None, nb=None):
    """matching.py:86-100 in float32: d2 = na + nb - 2 da.db, top-2, strict ratio, _unique."""
    if len(da) < 2 or len(db) < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    da = np.asarray(da, np.float32)
    db = np.asarray(db, np.float32)
    na = (da * da).sum(1) if na is None else na
    nb = (db * db).sum(1) if nb is None else nb
    d2 = na[:, None] + nb[None, :] - np.float32(2.0) * (da @ db.T)
    first, a, bmax = _two_smallest(d2)
    b1 = np.maximum(a, 0)
    b2 = np.maximum(bmax, 0)
    ok = np.sqrt(b1) < np.float32(ratio) * np.sqrt(b2)
    r = np.arange(len(da))
    return unique(r[ok].astype(np.int64), first[ok].astype(np.int64), b1[ok].astype(np.float32))


def predict(xy: np.ndarray, Hm: np.ndarray) -> np.ndarray:
    """cv2.perspectiveTransform of points (matching.py:115) in float64: w = (h6 x + h7 y) + h8;
    |w| > FLT_EPSILON ? ((h0 x + h1 y) + h2) * (1/w) : 0 (OpenCV's perspectiveTransform_)."""
    xy = np.asarray(xy, np.float64)
    Hm = np.asarray(Hm, np.float64).reshape(9)
    x, y = xy[:, 0], xy[:, 1]
    w = (x * Hm[6] + y * Hm[7]) + Hm[8]
    ok = np.abs(w) > np.finfo(np.float32).eps
    iw = np.where(ok, 1.0 / np.where(ok, w, 1.0), 0.0)
    px = ((x * Hm[0] + y * Hm[1]) + Hm[2]) * iw
    py = ((x * Hm[3] + y * Hm[4]) + Hm[5]) * iw
    return np.where(ok[:, None], np.c_[px, py], 0.0)


def guided_candidates(pred: np.ndarray, xy_b: np.ndarray, wh_b, radius: float, kc: int):
    """matching.py:115-131: queries predicted inside B (with a margin of `radius`), each with
    the kc spatially nearest B keypoints strictly within `radius` (kc <= 0: all of them).
    Spatial distances are squared float64 (dx*dx + dy*dy < radius*radius, cKDTree's p=2
    comparison); ties in distance go to the lower index.
    Returns (query index array, list of candidate index arrays)."""
    pred = np.asarray(pred, np.float64)
    xy_b = np.asarray(xy_b, np.float64)
    Wb, Hb = wh_b
    inside = np.nonzero((pred[:, 0] > -radius) & (pred[:, 0] < Wb + radius)
                        & (pred[:, 1] > -radius) & (pred[:, 1] < Hb + radius))[0]
    qs, cands = [], []
    if len(xy_b) == 0:
        return np.zeros(0, np.int64), []
    r2 = float(radius) * float(radius)
    for q in inside:
        dx = xy_b[:, 0] - pred[q, 0]
        dy = xy_b[:, 1] - pred[q, 1]
        d = dx * dx + dy * dy
        idx = np.nonzero(d < r2)[0]
        if len(idx) == 0:
            continue
        if kc > 0 and len(idx) > kc:
            o = np.lexsort((idx, d[idx]))[:kc]
            idx = np.sort(idx[o])
        qs.append(q)
        cands.append(idx.astype(np.int64))
    return np.asarray(qs, np.int64), cands


def guided_f32(da, db, qs, cands, ratio: float):
    """matching.py:132-152 in float32 on explicit candidate lists."""
    da = np.asarray(da, np.float32)
    db = np.asarray(db, np.float32)
    na = (da * da).sum(1)
    nb = (db * db).sum(1)
    out_q, out_t, out_d = [], [], []
    for q, c in zip(qs, cands):
        d2 = np.maximum(na[q] + nb[c] - np.float32(2.0) * (db[c] @ da[q]), np.float32(0))
        best, b1, b2 = _two_smallest(d2[None, :])
        b1, b2 = b1[0], b2[0]
        if np.isfinite(b1) and np.sqrt(b1) < np.float32(ratio) * np.sqrt(b2):
            out_q.append(q)
            out_t.append(c[best[0]])
            out_d.append(b1)
    return unique(np.asarray(out_q, np.int64), np.asarray(out_t, np.int64), np.asarray(out_d, np.float32))


def dist64(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """sum_k (a_k - b_k)^2 in float64, k in order (rows of a against rows of b)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    acc = np.zeros((a.shape[0], b.shape[0]), np.float64)
    for k in range(a.shape[1]):
        t = a[:, k][:, None] - b[:, k][None, :]
        acc += t * t
    return acc


def match_fp64(da, db, ratio: float, qs=None, cands=None, block: int = 1024):
    """The FP64 oracle (module docstring).  Brute force when qs/cands are None (needs >= 2 on
    each side, as matching.py:88-89), else on the given candidate lists.
    Returns (q int64, t int64, b1 float64) sorted by t."""
    out_q, out_t, out_d = [], [], []
    if qs is None:
        if len(da) < 2 or len(db) < 2:
            return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
        for s0 in range(0, len(da), block):
            d = dist64(da[s0:s0 + block], db)
            best, b1, b2 = _two_smallest(d)
            ok = np.sqrt(b1) < float(ratio) * np.sqrt(b2)
            r = np.nonzero(ok)[0]
            out_q.append(r + s0)
            out_t.append(best[r])
            out_d.append(b1[r])
    else:
        for q, c in zip(qs, cands):
            d = dist64(np.asarray(da)[q:q + 1], np.asarray(db)[c])
            best, b1, b2 = _two_smallest(d)
            if np.sqrt(b1[0]) < float(ratio) * np.sqrt(b2[0]):
                out_q.append(np.array([q]))
                out_t.append(np.array([c[best[0]]]))
                out_d.append(b1)
    if not out_q:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    q = np.concatenate(out_q).astype(np.int64)
    t = np.concatenate(out_t).astype(np.int64)
    d = np.concatenate(out_d).astype(np.float64)
    order = np.lexsort((q, d, t))
    t_s = t[order]
    first = _first_of_runs(t_s)
    keep = order[first]
    return q[keep], t[keep], d[keep]


def match_fp64_fast(da, db, ratio: float, block: int = 2048):
    """match_fp64 (brute force) for large sets, same definition and result.  A float64 BLAS
    product gives every distance by the expansion |a|^2 + |b|^2 - 2 a.b, whose error e is at
    most ~128 * 2^-53 * (|a|^2 + |b|^2 + 2|a||b|); per query, every target whose expanded
    distance is within tol >= 2e of the second smallest expanded distance is rescored exactly
    (sum_k (a_k - b_k)^2 in float64, k in order).  The exact two smallest distances always lie in
    that set: a target with exact distance <= the exact second smallest s2 has an expanded
    distance <= s2 + e, and the second smallest expanded distance is >= s2 - e.  Selection,
    ties and the one-query-per-target rule then follow match_fp64.
    Returns (q int64, t int64, b1 float64) sorted by t."""
    da = np.asarray(da, np.float64)
    db = np.asarray(db, np.float64)
    if len(da) < 2 or len(db) < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    na = (da * da).sum(1)
    nb = (db * db).sum(1)
    tol = 1e-9 * (na.max() + nb.max()) + 1e-9
    out_q, out_t, out_d = [], [], []
    for s0 in range(0, len(da), block):
        A = da[s0:s0 + block]
        approx = na[s0:s0 + block, None] + nb[None, :] - 2.0 * (A @ db.T)
        second = np.partition(approx, 1, axis=1)[:, 1]
        qq, tt = np.nonzero(approx <= (second + tol)[:, None])
        acc = np.zeros(len(qq))
        for k in range(da.shape[1]):
            t_ = A[qq, k] - db[tt, k]
            acc += t_ * t_
        order = np.lexsort((tt, acc, qq))          # per query: distance, then index
        qq, tt, acc = qq[order], tt[order], acc[order]
        starts = np.nonzero(np.r_[True, qq[1:] != qq[:-1]])[0]
        b1 = acc[starts]
        b2 = acc[starts + 1]                        # >= 2 candidates: the second smallest is in the set
        ok = np.sqrt(b1) < float(ratio) * np.sqrt(b2)
        out_q.append(qq[starts][ok] + s0)
        out_t.append(tt[starts][ok])
        out_d.append(b1[ok])
    q = np.concatenate(out_q).astype(np.int64)
    t = np.concatenate(out_t).astype(np.int64)
    d = np.concatenate(out_d)
    order = np.lexsort((q, d, t))
    t_s = t[order]
    first = _first_of_runs(t_s)
    keep = order[first]
    return q[keep], t[keep], d[keep]


def guided_fp64_fast(da, xy_a, db, xy_b, Hm, wh_b, radius: float, ratio: float):
    """Vectorised FP64 oracle of guided matching with kc = 0 (all candidates in the window):
    the same definition as guided_candidates + match_fp64, for large pairs.
    Returns (q int64, t int64) sorted by t."""
    from scipy.spatial import cKDTree
    pred = predict(xy_a, Hm)
    Wb, Hb = wh_b
    inside = np.nonzero((pred[:, 0] > -radius) & (pred[:, 0] < Wb + radius)
                        & (pred[:, 1] > -radius) & (pred[:, 1] < Hb + radius))[0]
    xy_b = np.asarray(xy_b, np.float64)
    if len(inside) == 0 or len(xy_b) == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    lists = cKDTree(xy_b).query_ball_point(pred[inside], radius * (1 + 1e-9) + 1e-9)
    lens = np.array([len(x) for x in lists])
    qq = np.repeat(inside, lens)
    tt = np.concatenate([np.asarray(x, np.int64) for x in lists]) if lens.sum() else np.zeros(0, np.int64)
    dx = xy_b[tt, 0] - pred[qq, 0]
    dy = xy_b[tt, 1] - pred[qq, 1]
    ok = dx * dx + dy * dy < float(radius) * float(radius)
    qq, tt = qq[ok], tt[ok]
    a = np.asarray(da, np.float64)
    b = np.asarray(db, np.float64)
    d = np.zeros(len(qq))
    for s0 in range(0, len(qq), 1 << 16):
        sl = slice(s0, s0 + (1 << 16))
        acc = np.zeros(len(qq[sl]))
        A = a[qq[sl]]
        Bt = b[tt[sl]]
        for k in range(a.shape[1]):
            t = A[:, k] - Bt[:, k]
            acc += t * t
        d[sl] = acc
    order = np.lexsort((tt, d, qq))
    qq, tt, d = qq[order], tt[order], d[order]
    first = np.r_[True, qq[1:] != qq[:-1]]
    starts = np.nonzero(first)[0]
    cnt = np.diff(np.r_[starts, len(qq)])
    b1 = d[starts]
    b2 = np.where(cnt > 1, d[np.minimum(starts + 1, len(d) - 1)], np.inf)
    acc_ = np.sqrt(b1) < float(ratio) * np.sqrt(b2)
    q = qq[starts][acc_]
    t = tt[starts][acc_]
    dd = b1[acc_]
    o = np.lexsort((q, dd, t))
    t_s = t[o]
    f = _first_of_runs(t_s)
    keep = o[f]
    return q[keep], t[keep]


def guided_f32_fast(da, xy_a, db, xy_b, Hm, wh_b, radius: float, ratio: float, chunk: int = 1024):
    """The reference's float32 guided matcher (matching.py:113-152) vectorised the way it is
    written there, with every candidate inside the window (kc = all, through
    cKDTree.query_ball_point on the strict radius): chunks of `chunk` queries, candidate lists
    padded to the chunk's longest, d = |a|^2 + |b|^2 - 2 b.a in float32 by one batched product,
    padding at +inf, the two smallest by argpartition, the float32 ratio test, then the
    one-query-per-target rule.  The CPU baseline of the guided matcher (numpy, one thread).
    Returns (q int64, t int64)."""
    from scipy.spatial import cKDTree
    pred = predict(xy_a, Hm)
    Wb, Hb = wh_b
    inside = np.nonzero((pred[:, 0] > -radius) & (pred[:, 0] < Wb + radius)
                        & (pred[:, 1] > -radius) & (pred[:, 1] < Hb + radius))[0]
    xy_b = np.asarray(xy_b, np.float64)
    if len(inside) == 0 or len(xy_b) == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    lists = cKDTree(xy_b).query_ball_point(pred[inside], radius * (1 + 1e-9) + 1e-9)
    a = np.asarray(da, np.float32)
    b = np.asarray(db, np.float32)
    na = (a * a).sum(1)
    nb = (b * b).sum(1)
    r2 = float(radius) * float(radius)
    qs, ts, ds = [], [], []
    for s0 in range(0, len(inside), chunk):
        qi = inside[s0:s0 + chunk]
        cand = []
        for q, lst in zip(qi, lists[s0:s0 + chunk]):
            c = np.asarray(lst, np.int64)
            d = xy_b[c] - pred[q]
            cand.append(c[(d * d).sum(1) < r2])
        width = max(1, max(len(c) for c in cand))
        idx = np.zeros((len(qi), width), np.int64)
        valid = np.zeros((len(qi), width), bool)
        for r, c in enumerate(cand):
            idx[r, :len(c)] = c
            valid[r, :len(c)] = True
        d2 = na[qi, None] + nb[idx] - np.float32(2.0) * np.matmul(b[idx], a[qi][:, :, None])[:, :, 0]
        d2 = np.where(valid, np.maximum(d2, np.float32(0)), np.float32(np.inf))
        two = np.argpartition(d2, 1, axis=1)[:, :2] if width > 1 else np.zeros((len(qi), 1), np.int64)
        rr = np.arange(len(qi))
        v0 = d2[rr, two[:, 0]]
        v1 = d2[rr, two[:, 1]] if width > 1 else np.full(len(qi), np.inf, np.float32)
        best = np.where(v0 <= v1, two[:, 0], two[:, 1] if width > 1 else two[:, 0])
        b1 = np.minimum(v0, v1)
        b2 = np.maximum(v0, v1)
        ok = np.isfinite(b1) & (np.sqrt(b1) < np.float32(ratio) * np.sqrt(b2))
        qs.append(qi[ok])
        ts.append(idx[rr[ok], best[ok]])
        ds.append(b1[ok])
    return unique(np.concatenate(qs), np.concatenate(ts), np.concatenate(ds))
