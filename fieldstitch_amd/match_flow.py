"""match_pair for many pairs at once (the per-pair process pool of pipeline.py:535-548 over
matching.py:205-253, as one batched launch sequence per stage).

Every pair goes through the reference's decision chain; what changes is that each stage runs
for all the pairs that reach it in one batch:

  1. coarse      the n_coarse strongest features of both photos, exhaustive ratio matching at
                 0.85 (matching.py:219-220): one dense PairMatcher batch;
  2. prior       the ground-plane prior filter (matching.py:221-225): one fsx_prior_keep_batch;
  3. homography  RANSAC on pairs with >= 12 coarse matches, threshold 0.012 max(W_i, H_i)
                 (matching.py:226): one fsx_ransac_batch_adaptive, a seed per pair;
  4. guided      pairs whose homography has >= 10 inliers: guided matching in a window of
                 guided_radius max(W_j, H_j) around the homography's prediction, the guided_k
                 nearest, ratio 0.85 (matching.py:227-229): one PairMatcher batch per window
                 size, optionally pruned by the learned match filter;
  5. verify      pairs with >= min_inliers guided matches: fundamental-matrix RANSAC at
                 ransac_thr; success needs max(min_inliers, 40) inliers (matching.py:230-233);
  6. fallback    the rest: exhaustive ratio matching at `ratio` (the exact set FLANN
                 approximates, matching.py:236), prior filter, >= 20 survivors, verification
                 with success at min_inliers (matching.py:236-246);
  7. similarity  RANSAC of a 2-point similarity at 0.02 max(W_i, H_i), then its least-squares
                 fit on the inliers (matching.py:194-202).

The three RANSAC stages stop as cv2's estimators do: a pair draws hypotheses in rounds of 256
until its best model's inlier ratio w makes (1 - w^K)^(hypotheses drawn) <= 1 - confidence,
at most the call's iteration cap (confidence / cap of each call in H_CONF, F_CONF, S_CONF).

The match lists stay on the device from one stage to the next (packed per pair; the
matcher's output packed by fsx_match_pack, lists chosen, cut to the prior's survivors or to
RANSAC inliers, and gathered with their positions by fsx_gather_lists): the host handles only
per-pair counts, models and fits, and copies the verified lists once at the end.

Results are the reference's tuples: (i, j, q int32, t int32, sim, n_inliers), or
(i, j, None, None, None, n) when matching fails.  A stage that raises sends its pairs through
the chain one by one, so a failing pair reports ("error", message) in the sim slot
(matching.py:249-253) without taking the others down.  Each pair's RANSAC draws come from a
seed of its own (i, j), so a pair's tuple does not depend on the batch it ran in.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
import traceback

import numpy as np

from . import _lib

# each estimator's iteration cap and confidence, as the reference calls cv2 (matching.py:226
# findHomography with its defaults, :180 findFundamentalMat(..., 0.999, 10000), :196
# estimateAffinePartial2D with its defaults); the device RANSAC stops as they do, once
# the best model's inlier ratio w gives (1 - w^K)^(hypotheses drawn) <= 1 - confidence
H_ITERS, H_CONF = 2000, 0.995
F_ITERS, F_CONF = 10000, 0.999
S_ITERS, S_CONF = 2000, 0.99
_MODEL = {"homography": 0, "fundamental": 1, "similarity": 2}


def pair_seed(i, j, salt):
    """RANSAC seed of the image pair (i, j) and stage `salt`."""
    return (0x5EED0000 ^ (int(i) * 0x9E3779B1) ^ (int(j) << 24) ^ salt) & 0xFFFFFFFFFFFFFFFF


class FlowImage:
    """One photo's device-resident features: RootSIFT uint8 descriptors (n, 128) in response
    order, undistorted full-resolution positions (n, 2) float64, the photo size and, for the
    learned filter, keypoint size and angle (float32, may be None)."""

    def __init__(self, desc, und, W, H, size=None, angle=None):
        self.desc = desc
        self.und = und
        self.W = int(W)
        self.H = int(H)
        self.size = size
        self.angle = angle
        self.n = int(desc.shape[0])
        self._und_host = None

    def und_host(self):
        if self._und_host is None:
            self._und_host = self.und.cpu().numpy()
        return self._und_host


class FlowConfig:
    """The matching context's parameters (pipeline.py:478-482): per-image W / H and attitude
    flags, the prior's cameras, and the thresholds."""

    def __init__(self, ctx):
        self.W = list(ctx["W"])
        self.H = list(ctx["H"])
        self.has_rot = list(ctx.get("has_rot", [False] * len(self.W)))
        self.use_prior = bool(ctx.get("use_prior", True))
        self.prior_tol = float(ctx.get("prior_tol", 0.18))
        self.n_coarse = int(ctx.get("n_coarse", 1500))
        self.guided = bool(ctx.get("guided", True))
        self.guided_radius = float(ctx.get("guided_radius", 0.03))
        self.guided_k = int(ctx.get("guided_k", 24))
        self.min_inliers = int(ctx["min_inliers"])
        self.ransac_thr = float(ctx["ransac_thr"])
        self.f_iters = int(ctx.get("ransac_iters", F_ITERS))
        self.ratio = float(ctx.get("ratio", 0.8))
        self._cams = None
        self._halves = None
        if all(k in ctx for k in ("K", "R", "C", "zg")):
            self._cams = (ctx["K"], ctx["R"], ctx["C"], ctx["zg"])

    def prior_on(self, i, j):
        return bool(self.has_rot[i] and self.has_rot[j] and self.use_prior and self._cams is not None)

    def _cam_rows(self):
        """Per camera: the 22 float64 of its source half of the prior record (K^-1, R, C, z_g)
        and the 21 of its target half (R, K, C), computed once."""
        if self._halves is None:
            K, R, Cc, zg = self._cams
            n = len(K)
            src = np.zeros((n, 22), np.float64)
            dst = np.zeros((n, 21), np.float64)
            for c in range(n):
                Kc = np.asarray(K[c], np.float64)
                src[c] = np.concatenate([np.linalg.inv(Kc).ravel(), np.asarray(R[c], np.float64).ravel(),
                                         np.asarray(Cc[c], np.float64).ravel(), [float(zg[c])]])
                dst[c] = np.concatenate([np.asarray(R[c], np.float64).ravel(), Kc.ravel(),
                                         np.asarray(Cc[c], np.float64).ravel()])
            self._halves = (src, dst)
        return self._halves

    def cams(self, i, j):
        """The 43 float64 of the ground-plane prior from camera i into camera j."""
        src, dst = self._cam_rows()
        return np.concatenate([src[i], dst[j]])

    def side(self, i):
        return max(self.W[i], self.H[i])


# ---- batched primitives ---------------------------------------------------------------------

def _torch():
    import torch
    return torch


def _concat_rows(rows):
    lens = np.array([len(r) for r in rows], np.int64)
    off = np.r_[0, np.cumsum(lens)].astype(np.int64)
    data = np.concatenate(rows) if off[-1] else np.zeros((0, 2))
    return off, np.ascontiguousarray(data, np.float64).reshape(-1, 2)


class _Rows:
    """Features q of the photo whose rows start at base in a Positions array."""
    __slots__ = ("base", "q")

    def __init__(self, base, q):
        self.base = base
        self.q = q

    def __len__(self):
        return len(self.q)


class Positions:
    """All photos' undistorted positions in one device float64 array: photo i's rows start at
    base[i], so a pair's matched rows are gathered by index (fsx_gather_xy) instead of being
    indexed, packed and uploaded on the host."""

    def __init__(self, arrays):
        torch = _torch()
        self.base = {}
        parts = []
        o = 0
        for i, a in arrays.items():
            t = a if isinstance(a, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(a, np.float64))
            t = t.to("cuda", torch.float64).reshape(-1, 2)
            self.base[i] = o
            o += int(t.shape[0])
            parts.append(t)
        if o >= 2 ** 31:
            raise ValueError("%d positions in all; at most 2^31 - 1" % o)
        self.n = o
        self.xy = torch.cat(parts).contiguous() if o else torch.zeros((1, 2), dtype=torch.float64, device="cuda")

    def base_of(self, photos):
        """Row bases of photos (int64; a photo not held gives a base no row reaches)."""
        return np.array([self.base.get(int(i), -(1 << 62)) for i in photos], np.int64)

    def rows(self, i, q):
        """Photo i's features q as a row reference for gather (global row numbers base[i] + q;
        the addition is done once for a whole list, in gather)."""
        return _Rows(self.base[i], q)

    def gather(self, idx_list):
        """[row references of rows(), or (m_k,) global row numbers] -> (int64 offsets, device
        float64 (sum m_k, 2) rows)."""
        torch = _torch()
        lens = np.fromiter((len(x.q) if isinstance(x, _Rows) else len(x) for x in idx_list), np.int64,
                           count=len(idx_list))
        off = np.zeros(len(idx_list) + 1, np.int64)
        np.cumsum(lens, out=off[1:])
        tot = int(off[-1])
        out = torch.empty((max(tot, 1), 2), dtype=torch.float64, device="cuda")
        if tot:
            base = np.fromiter((x.base if isinstance(x, _Rows) else 0 for x in idx_list), np.int64, count=len(idx_list))
            rows = np.concatenate([x.q if isinstance(x, _Rows) else x for x in idx_list]).astype(np.int64, copy=False)
            idx = torch.from_numpy((rows + np.repeat(base, lens)).astype(np.int32)).to("cuda")
            _lib.call("fsx_gather_xy", tot, C.c_void_p(idx.data_ptr()), C.c_void_p(self.xy.data_ptr()), self.n,
                      C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
        return off, out


class _Masks(list):
    """Per-pair inlier masks (views of one flat bool array; .flat and .off keep the whole)."""

    def __init__(self, flat, off):
        super().__init__(flat[off[k]:off[k + 1]] for k in range(len(off) - 1))
        self.flat = flat[:int(off[-1])]
        self.off = off


def _ransac_call(model, n, d_off, pa, pb, n_hyp, d_thr, d_seed, inl, M, cnt, conf):
    """fsx_ransac_batch (conf None: n_hyp hypotheses per pair) or fsx_ransac_batch_adaptive
    (at most n_hyp, stopping at confidence conf)."""
    Pv = C.c_void_p
    head = (_MODEL[model], n, Pv(d_off.data_ptr()), Pv(pa.data_ptr()), Pv(pb.data_ptr()), int(n_hyp))
    tail = (Pv(d_thr.data_ptr()), Pv(d_seed.data_ptr()), Pv(inl.data_ptr()), Pv(M.data_ptr()), Pv(cnt.data_ptr()))
    if conf is None:
        _lib.call("fsx_ransac_batch", *head, *tail, Pv(_lib.stream_handle()))
    else:
        _lib.call("fsx_ransac_batch_adaptive", *head, float(conf), *tail, None, Pv(_lib.stream_handle()))


def _ransac_dev(model, off, d_a, d_b, thr, seeds, n_hyp, conf=None):
    """fsx_ransac_batch (or, with conf, the adaptive form) on device rows (pair k's rows off[k]
    .. off[k+1] of d_a / d_b).  Returns (inlier masks, models (n, 9), counts, the device inlier
    bytes)."""
    torch = _torch()
    n = len(off) - 1
    dev = "cuda"
    tot = max(int(off[-1]), 1)
    d_off = torch.from_numpy(off).to(dev)
    d_thr = torch.from_numpy(np.asarray(thr, np.float64).reshape(n)).to(dev)
    d_seed = torch.from_numpy(np.asarray([s & 0x7FFFFFFFFFFFFFFF if s < (1 << 63) else s - (1 << 64) for s in seeds],
                                         np.int64)).to(dev)
    inl = torch.zeros(tot, dtype=torch.uint8, device=dev)
    M = torch.zeros((n, 9), dtype=torch.float64, device=dev)
    cnt = torch.zeros(n, dtype=torch.int32, device=dev)
    _ransac_call(model, n, d_off, d_a, d_b, n_hyp, d_thr, d_seed, inl, M, cnt, conf)
    flags = inl.cpu().numpy().view(bool)
    return (_Masks(flags, off), M.cpu().numpy(), cnt.cpu().numpy().astype(np.int64), (d_off, inl))


def ransac_batch(model, pa_list, pb_list, thr, seeds, n_hyp, conf=None):
    """RANSAC of `model` for every pair, each with its own threshold and seed (one launch; with
    conf, adaptive stopping at that confidence, n_hyp the cap), on host (n_k, 2) position
    arrays.  Returns (inlier masks, models (n, 9), counts)."""
    torch = _torch()
    if not pa_list:
        return [], np.zeros((0, 9)), np.zeros(0, np.int64)
    off, A = _concat_rows(pa_list)
    _, B = _concat_rows(pb_list)
    d_a = torch.from_numpy(A if off[-1] else np.zeros((1, 2))).to("cuda")
    d_b = torch.from_numpy(B if off[-1] else np.zeros((1, 2))).to("cuda")
    flags, M, cnt, _ = _ransac_dev(model, off, d_a, d_b, thr, seeds, n_hyp, conf)
    return flags, M, cnt


def ransac_rows(model, P, ia_list, ib_list, thr, seeds, n_hyp, fit=False, conf=None):
    """ransac_batch on the rows ia_list[k] / ib_list[k] of the Positions P, gathered on the
    device.  With fit=True also the least-squares similarity of each pair's inliers
    (fsx_similarity_fit_batch): a (n, 4) array of (angle, scale, tx, ty), NaN rows without a
    fit.  Returns (inlier masks, models, counts[, fits])."""
    torch = _torch()
    if not ia_list:
        empty = ([], np.zeros((0, 9)), np.zeros(0, np.int64))
        return empty + (np.zeros((0, 4)),) if fit else empty
    off, d_a = P.gather(ia_list)
    _, d_b = P.gather(ib_list)
    flags, M, cnt, (d_off, inl) = _ransac_dev(model, off, d_a, d_b, thr, seeds, n_hyp, conf)
    if not fit:
        return flags, M, cnt
    n = len(ia_list)
    out = torch.empty((n, 4), dtype=torch.float64, device="cuda")
    Pv = C.c_void_p
    _lib.call("fsx_similarity_fit_batch", n, Pv(d_off.data_ptr()), Pv(d_a.data_ptr()), Pv(d_b.data_ptr()),
              Pv(inl.data_ptr()), Pv(out.data_ptr()), Pv(_lib.stream_handle()))
    return flags, M, cnt, out.cpu().numpy()


def similarity_fit(a, b):
    """Least-squares similarity b ~ s R(theta) a + t of matched points (closed form on the
    centred sets): (theta rad, s, tx, ty), or None for fewer than 2 points or points that
    coincide.  The host form of fsx_similarity_fit_batch (the chain uses the device one)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    if len(a) < 2:
        return None
    ma, mb = a.mean(0), b.mean(0)
    ac, bc = a - ma, b - mb
    den = float((ac * ac).sum())
    if not den > 1e-12 * len(a):        # all points within ~1e-6 px of their mean: no fit
        return None
    c = float((ac[:, 0] * bc[:, 0] + ac[:, 1] * bc[:, 1]).sum()) / den
    s = float((ac[:, 0] * bc[:, 1] - ac[:, 1] * bc[:, 0]).sum()) / den
    tx = mb[0] - (c * ma[0] - s * ma[1])
    ty = mb[1] - (s * ma[0] + c * ma[1])
    return (math.atan2(s, c), math.hypot(c, s), float(tx), float(ty))


# ---- the chain ------------------------------------------------------------------------------

def _error_tuple(i, j, e):
    tb = traceback.extract_tb(e.__traceback__)
    where = tb[-1] if tb else None
    loc = " (%s line %d)" % (os.path.basename(where.filename), where.lineno) if where else ""
    return (i, j, None, None, ("error", "%s: %s%s" % (type(e).__name__, e, loc)), 0)


def run_pairs(tasks, images, cfg, form="mfma", filt=None, stats=None):
    """The match_pair chain for every (i, j) of `tasks` over `images` (index -> FlowImage).
    Returns the tuples in task order.  filt = (net, tau) prunes the guided stage's candidates
    with the learned filter.  stats (dict, optional) receives per-stage pair counts and host
    wall times (stage_ms)."""
    tasks = [(int(i), int(j)) for i, j in tasks]
    try:
        return _chain(tasks, images, cfg, form, filt, stats)
    except Exception as e:  # noqa: BLE001 - isolate the failing pair(s), as the per-pair pool does
        if len(tasks) == 1:
            return [_error_tuple(tasks[0][0], tasks[0][1], e)]
        return [run_pairs([t], images, cfg, form, filt)[0] for t in tasks]


# ---- the chain's match lists on the device --------------------------------------------------

class _DevLists:
    """Match lists on the device: list k holds the (q, t) feature pairs of photos a[k] -> b[k] in
    rows start[k] .. start[k] + cnt[k] of the int32 tensors q and t (host start / cnt int64)."""

    def __init__(self, q, t, start, cnt, a, b):
        self.q, self.t = q, t
        self.start = np.asarray(start, np.int64)
        self.cnt = np.asarray(cnt, np.int64)
        self.a = np.asarray(a, np.int64)
        self.b = np.asarray(b, np.int64)

    def __len__(self):
        return len(self.cnt)

    def off(self):
        """Row offsets (n + 1) of lists packed in order (as _move writes them)."""
        o = np.zeros(len(self.cnt) + 1, np.int64)
        np.cumsum(self.cnt, out=o[1:])
        return o


def _to_host(*xs):
    """Device tensors -> numpy arrays through pinned host buffers: one DMA each on the current
    stream, then one synchronisation (a pageable copy of tens of MB goes through many small
    staged copies at a few GB/s)."""
    torch = _torch()
    hs = [torch.empty(x.shape, dtype=x.dtype, pin_memory=True) for x in xs]
    for h, x in zip(hs, xs):
        h.copy_(x, non_blocking=True)
    torch.cuda.current_stream().synchronize()
    return [h.numpy() for h in hs]


def _dev_i64(*arrays):
    """Host int64 arrays of one length -> one device (k, n) int64 tensor (a single upload)."""
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(np.stack([np.asarray(a, np.int64) for a in arrays]))).to("cuda")


def _move(sources, n_out, P=None):
    """A new _DevLists of n_out lists packed in order: for each source (L, src, dst, keep),
    list dst[i] = L's list src[i], only its rows whose byte in keep (device uint8 over L's rows,
    or None: all rows) is set.  With the Positions P also the rows' positions (pa, pb) from one
    gather.  Returns (lists, pa, pb)."""
    torch = _torch()
    cnt = np.zeros(n_out, np.int64)
    a = np.zeros(n_out, np.int64)
    b = np.zeros(n_out, np.int64)
    plans = []
    for L, src, dst, keep in sources:
        src = np.asarray(src, np.int64)
        dst = np.asarray(dst, np.int64)
        if len(src) == 0:
            continue
        d_src = _dev_i64(L.start[src])
        d_cnt = torch.from_numpy(L.cnt[src].astype(np.int32)).to("cuda")
        if keep is None:
            c = L.cnt[src]
        else:                   # counting pass: the rows each list keeps
            kept = torch.zeros(len(src), dtype=torch.int32, device="cuda")
            _lib.call("fsx_gather_lists", len(src), C.c_void_p(d_src.data_ptr()), C.c_void_p(d_cnt.data_ptr()),
                      C.c_void_p(d_src.data_ptr()), 0, None, None, C.c_void_p(L.q.data_ptr()),
                      C.c_void_p(L.t.data_ptr()), C.c_void_p(keep.data_ptr()), None, 0, None, None, None, None,
                      C.c_void_p(kept.data_ptr()), C.c_void_p(_lib.stream_handle()))
            c = kept.cpu().numpy().astype(np.int64)
        cnt[dst] = c
        a[dst] = L.a[src]
        b[dst] = L.b[src]
        plans.append((L, src, dst, keep, d_src, d_cnt))
    start = np.zeros(n_out, np.int64)
    if n_out:
        start[1:] = np.cumsum(cnt)[:-1]
    tot = int(cnt.sum())
    q = torch.empty(max(tot, 1), dtype=torch.int32, device="cuda")
    t = torch.empty(max(tot, 1), dtype=torch.int32, device="cuda")
    pa = pb = None
    if P is not None:
        pa = torch.empty((max(tot, 1), 2), dtype=torch.float64, device="cuda")
        pb = torch.empty((max(tot, 1), 2), dtype=torch.float64, device="cuda")
    Pv = C.c_void_p
    for L, src, dst, keep, d_src, d_cnt in plans:
        if P is not None:
            d_aux = _dev_i64(start[dst], P.base_of(L.a[src]), P.base_of(L.b[src]))
            d_dst, d_ba, d_bb = (Pv(d_aux[r].data_ptr()) for r in range(3))
        else:
            d_aux = _dev_i64(start[dst])
            d_dst, d_ba, d_bb = Pv(d_aux[0].data_ptr()), None, None
        _lib.call("fsx_gather_lists", len(src), Pv(d_src.data_ptr()), Pv(d_cnt.data_ptr()), d_dst, tot, d_ba, d_bb,
                  Pv(L.q.data_ptr()), Pv(L.t.data_ptr()), None if keep is None else Pv(keep.data_ptr()),
                  Pv(P.xy.data_ptr()) if P is not None else None, P.n if P is not None else 0, Pv(q.data_ptr()),
                  Pv(t.data_ptr()), Pv(pa.data_ptr()) if pa is not None else None,
                  Pv(pb.data_ptr()) if pb is not None else None, None, Pv(_lib.stream_handle()))
    return _DevLists(q, t, start, cnt, a, b), pa, pb


def _lists_of_matcher(pm, pairs):
    """The last run of PairMatcher pm (pairs [(i, j)] in its order) as _DevLists."""
    cnt, off, q, t = pm.results_dev()
    return _DevLists(q, t, off[:-1], cnt, [i for i, _ in pairs], [j for _, j in pairs])


def _dense_dev(images, pairs, ratio, prefix=None, form="mfma"):
    """_dense_matches with the lists left on the device."""
    from .matching import PairMatcher
    torch = _torch()
    if not pairs:
        z = torch.empty(1, dtype=torch.int32, device="cuda")
        return _DevLists(z, z, [], [], [], [])
    used = sorted({i for p in pairs for i in p})
    slot = {i: k for k, i in enumerate(used)}
    zeros = {}
    sets = []
    for i in used:
        d = images[i].desc if prefix is None else images[i].desc[:prefix]
        n = int(d.shape[0])
        if n not in zeros:
            zeros[n] = torch.zeros((n, 2), dtype=torch.float64, device=d.device)
        sets.append((d, zeros[n], (1, 1)))
    pm = PairMatcher(sets, [(slot[i], slot[j]) for i, j in pairs], None, radius=1.0, ratio=ratio)
    pm.run(form)
    L = _lists_of_matcher(pm, pairs)
    # the reference returns nothing when either side has fewer than 2 features (matching.py:88)
    for k, (i, j) in enumerate(pairs):
        if min(sets[slot[i]][0].shape[0], sets[slot[j]][0].shape[0]) < 2:
            L.cnt[k] = 0
    return L


def _guided_dev(images, pairs, Hms, radii, kc, ratio, form="mfma", filt=None):
    """_guided_matches with the lists left on the device, in the order of `pairs`."""
    from .matching import PairMatcher
    torch = _torch()
    if not pairs:
        z = torch.empty(1, dtype=torch.int32, device="cuda")
        return _DevLists(z, z, [], [], [], [])
    parts = []
    groups = {}
    for k, r in enumerate(radii):
        groups.setdefault(float(r), []).append(k)
    for r, ks in groups.items():
        used = sorted({i for k in ks for i in pairs[k]})
        slot = {i: s for s, i in enumerate(used)}
        sets = [(images[i].desc, images[i].und, (images[i].W, images[i].H)) for i in used]
        kcs = [0 if kc >= images[pairs[k][1]].n else kc for k in ks]
        # kc >= n_b means every candidate in the window (cKDTree's k clipped to n_b)
        for kc_val in sorted(set(kcs)):
            ks2 = [k for k, v in zip(ks, kcs) if v == kc_val]
            sub2 = [(slot[pairs[k][0]], slot[pairs[k][1]]) for k in ks2]
            pm = PairMatcher(sets, sub2, [Hms[k] for k in ks2], radius=r, kc=kc_val, ratio=ratio)
            if filt is None:
                pm.run(form)
            else:
                net, tau = filt
                pm.run_filtered(form, net, tau, [images[i].size for i in used], [images[i].angle for i in used])
            parts.append((_lists_of_matcher(pm, [pairs[k] for k in ks2]), ks2))
    if len(parts) == 1 and parts[0][1] == list(range(len(pairs))):
        return parts[0][0]
    L, _, _ = _move([(Lp, np.arange(len(ks2)), ks2, None) for Lp, ks2 in parts], len(pairs))
    return L


def _prior_dev(cfg, P, L, sel):
    """L with lists sel (indices; prior on, non-empty) cut to the matches that pass the
    ground-plane prior (matching.py:221-225): one gather, one fsx_prior_keep_batch, one
    compaction."""
    torch = _torch()
    if not len(sel):
        return L
    sel = np.asarray(sel, np.int64)
    G, pa, pb = _move([(L, sel, np.arange(len(sel)), None)], len(sel), P)
    off = G.off()
    cams = np.zeros((len(sel), 43), np.float64)
    src, dst = cfg._cam_rows()
    cams[:, :22] = src[G.a]
    cams[:, 22:] = dst[G.b]
    tol = cfg.prior_tol * np.maximum(np.asarray(cfg.W, np.float64)[G.b], np.asarray(cfg.H, np.float64)[G.b])
    keep = torch.empty(max(int(off[-1]), 1), dtype=torch.uint8, device="cuda")
    d_off = torch.from_numpy(off).to("cuda")
    d_cams = torch.from_numpy(cams).to("cuda")
    d_tol = torch.from_numpy(tol).to("cuda")
    Pv = C.c_void_p
    _lib.call("fsx_prior_keep_batch", len(sel), Pv(d_off.data_ptr()), Pv(pa.data_ptr()), Pv(pb.data_ptr()),
              Pv(d_cams.data_ptr()), Pv(d_tol.data_ptr()), int(G.cnt.max()), Pv(keep.data_ptr()),
              Pv(_lib.stream_handle()))
    rest = np.setdiff1d(np.arange(len(L)), sel)
    out, _, _ = _move([(L, rest, rest, None), (G, np.arange(len(sel)), sel, keep)], len(L))
    return out


def _ransac_lists(model, off, pa, pb, thr, seeds, n_hyp, want_model=False, fit=False, conf=None):
    """fsx_ransac_batch (with conf: fsx_ransac_batch_adaptive) on lists packed by _move (rows
    off[k] .. off[k+1] of pa / pb), flags left on the device.  Returns (device inlier bytes,
    counts, models (n, 9) or None, similarity fits (n, 4) or None)."""
    torch = _torch()
    n = len(off) - 1
    tot = max(int(off[-1]), 1)
    d_off = torch.from_numpy(np.asarray(off, np.int64)).to("cuda")
    d_thr = torch.from_numpy(np.asarray(thr, np.float64).reshape(n)).to("cuda")
    d_seed = torch.from_numpy(np.asarray([s & 0x7FFFFFFFFFFFFFFF if s < (1 << 63) else s - (1 << 64) for s in seeds],
                                         np.int64)).to("cuda")
    inl = torch.zeros(tot, dtype=torch.uint8, device="cuda")
    M = torch.zeros((n, 9), dtype=torch.float64, device="cuda")
    cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    Pv = C.c_void_p
    _ransac_call(model, n, d_off, pa, pb, n_hyp, d_thr, d_seed, inl, M, cnt, conf)
    fits = None
    if fit:
        fits = torch.empty((n, 4), dtype=torch.float64, device="cuda")
        _lib.call("fsx_similarity_fit_batch", n, Pv(d_off.data_ptr()), Pv(pa.data_ptr()), Pv(pb.data_ptr()),
                  Pv(inl.data_ptr()), Pv(fits.data_ptr()), Pv(_lib.stream_handle()))
    return (inl, cnt.cpu().numpy().astype(np.int64), M.cpu().numpy() if want_model else None,
            fits.cpu().numpy() if fit else None)


def _verify_dev(cfg, tasks, keys, P, L, sel):
    """_verify on lists sel of L (keys: their task indices): F RANSAC of the lists with >= 8
    matches, then each list cut to its inliers on the device.  Returns (the inlier lists of the
    verified pairs, [(key, list index, n_in)] for them, {key: n_in} for every list of sel)."""
    sel = [n for n in sel if L.cnt[n] >= 8]
    n_in = {}
    if not sel:
        return None, [], n_in
    V, pa, pb = _move([(L, sel, np.arange(len(sel)), None)], len(sel), P)
    inl, cnt, _, _ = _ransac_lists("fundamental", V.off(), pa, pb, [cfg.ransac_thr] * len(sel),
                                   [pair_seed(*tasks[keys[n]], 0xF) for n in sel], cfg.f_iters, conf=F_CONF)
    ok = [m for m in range(len(sel)) if cnt[m] > 0]
    I, _, _ = _move([(V, ok, np.arange(len(ok)), inl)], len(ok))
    if not np.array_equal(I.cnt, cnt[ok]):
        raise RuntimeError("inlier flags disagree with the RANSAC counts")
    done = []
    for r, m in enumerate(ok):
        n_in[keys[sel[m]]] = int(I.cnt[r])
        done.append((keys[sel[m]], r, int(I.cnt[r])))
    return I, done, n_in


def _chain(tasks, images, cfg, form, filt, stats):
    n_task = len(tasks)
    out = [None] * n_task
    live = []
    for k, (i, j) in enumerate(tasks):
        if images[i].n < 30 or images[j].n < 30:
            out[k] = (i, j, None, None, None, 0)
        else:
            live.append(k)
    marks = [("start", time.perf_counter())]
    P = Positions({i: images[i].und for i in dict.fromkeys(i for k in live for i in tasks[k])})
    result = {}            # task index -> (inlier lists, list index, n_in)
    marks.append(("positions", time.perf_counter()))

    # 1-2. coarse matches of the strongest features, prior filter
    n1 = cfg.n_coarse
    coarse = [k for k in live if cfg.guided and images[tasks[k][0]].n > n1 // 2]
    Cm = _dense_dev(images, [tasks[k] for k in coarse], 0.85, prefix=n1, form=form)
    Cm = _prior_dev(cfg, P, Cm, [n for n, k in enumerate(coarse) if cfg.prior_on(*tasks[k]) and Cm.cnt[n] > 0])
    marks.append(("coarse_and_prior", time.perf_counter()))

    # 3. homography RANSAC on the coarse set
    hsel = [n for n in range(len(coarse)) if Cm.cnt[n] >= 12]
    hk = [coarse[n] for n in hsel]
    Hm = {}
    gk = []
    if hsel:
        Hl, pa, pb = _move([(Cm, hsel, np.arange(len(hsel)), None)], len(hsel), P)
        _, hcnt, hM, _ = _ransac_lists("homography", Hl.off(), pa, pb, [0.012 * cfg.side(tasks[k][0]) for k in hk],
                                       [pair_seed(*tasks[k], 0x1) for k in hk], H_ITERS, want_model=True,
                                       conf=H_CONF)
        gk = [k for k, c in zip(hk, hcnt.tolist()) if c >= 10]          # the winner's count = its flags' sum
        Hm = {k: hM[n].reshape(3, 3) for n, k in enumerate(hk)}
    marks.append(("homography", time.perf_counter()))

    # 4. guided matching under the homography
    G = _guided_dev(images, [tasks[k] for k in gk], [Hm[k] for k in gk],
                    [cfg.guided_radius * cfg.side(tasks[k][1]) for k in gk], cfg.guided_k, 0.85, form, filt)
    vsel = [n for n in range(len(gk)) if G.cnt[n] >= cfg.min_inliers]
    marks.append(("guided", time.perf_counter()))

    # 5. fundamental-matrix verification of the guided matches
    need = max(cfg.min_inliers, 40)
    I1, done, _ = _verify_dev(cfg, tasks, gk, P, G, vsel)
    for k, r, n_in in done:
        if n_in >= need:
            result[k] = (I1, r, n_in)
    marks.append(("verify", time.perf_counter()))

    # 6. fallback: exhaustive matching, prior filter, verification
    fb = [k for k in live if k not in result]
    Fm = _dense_dev(images, [tasks[k] for k in fb], cfg.ratio, form=form)
    Fm = _prior_dev(cfg, P, Fm, [n for n, k in enumerate(fb) if cfg.prior_on(*tasks[k]) and Fm.cnt[n] > 0])
    enough = []
    for n, k in enumerate(fb):
        if Fm.cnt[n] < 20:
            out[k] = (tasks[k][0], tasks[k][1], None, None, None, 0)
        else:
            enough.append(n)
    I2, done2, n_in2 = _verify_dev(cfg, tasks, fb, P, Fm, enough)
    for n in enough:
        k = fb[n]
        if k not in n_in2 or n_in2[k] < cfg.min_inliers:
            out[k] = (tasks[k][0], tasks[k][1], None, None, None, n_in2.get(k, 0))
    for k, r, n_in in done2:
        if n_in >= cfg.min_inliers:
            result[k] = (I2, r, n_in)

    marks.append(("fallback", time.perf_counter()))
    # 7. similarity of every verified set: the lists in task order on the device, one RANSAC
    # launch with its fits, then one copy of all verified lists to the host
    rk = sorted(result)
    if rk:
        srcs = []
        for Lx in {id(result[k][0]): result[k][0] for k in rk}.values():
            ks = [n for n, k in enumerate(rk) if result[k][0] is Lx]
            srcs.append((Lx, [result[rk[n]][1] for n in ks], ks, None))
        R, pa, pb = _move(srcs, len(rk), P)
        _, scnt, _, fits = _ransac_lists("similarity", R.off(), pa, pb, [0.02 * cfg.side(tasks[k][0]) for k in rk],
                                         [pair_seed(*tasks[k], 0x5) for k in rk], S_ITERS, fit=True, conf=S_CONF)
        good = ((scnt >= 2) & np.isfinite(fits).all(1)).tolist()
        tot = int(R.cnt.sum())
        q32, t32 = _to_host(R.q[:tot], R.t[:tot])
        o = R.off().tolist()
        for n, (k, g, fit) in enumerate(zip(rk, good, fits.tolist())):
            out[k] = (tasks[k][0], tasks[k][1], q32[o[n]:o[n + 1]], t32[o[n]:o[n + 1]], tuple(fit) if g else None,
                      result[k][2])
    marks.append(("similarity", time.perf_counter()))
    if stats is not None:
        stats.update({"pairs": n_task, "coarse": len(coarse), "homography": len(hk), "guided": len(gk),
                      "guided_verified": len(vsel), "fallback": len(fb), "matched": len(rk),
                      "stage_ms": {b[0]: round(1e3 * (b[1] - a[1]), 3) for a, b in zip(marks, marks[1:])}})
    return out


def _verify(cfg, tasks, P, items):
    """matching.py:175-191 for items [(task index, (q, t))]: fundamental-matrix RANSAC at
    ransac_thr (fewer than 8 matches: no model) -> [(q, t of the inliers, count) or
    (None, None, 0)]."""
    res = [(None, None, 0)] * len(items)
    ok = [n for n, (_, (q, _)) in enumerate(items) if len(q) >= 8]
    flags, _, cnt = ransac_rows("fundamental", P,
                                [P.rows(tasks[items[n][0]][0], items[n][1][0]) for n in ok],
                                [P.rows(tasks[items[n][0]][1], items[n][1][1]) for n in ok],
                                [cfg.ransac_thr] * len(ok), [pair_seed(*tasks[items[n][0]], 0xF) for n in ok],
                                cfg.f_iters, conf=F_CONF)
    if not ok:
        return res
    # the inliers of all pairs by one mask over the concatenated matches, split per pair
    qf = np.concatenate([items[n][1][0] for n in ok])[flags.flat]
    tf = np.concatenate([items[n][1][1] for n in ok])[flags.flat]
    n_f = np.add.reduceat(flags.flat.astype(np.int64), flags.off[:-1]).tolist()    # every pair has >= 8 rows
    ends = np.cumsum(n_f).tolist()
    for n, c, e, m in zip(ok, cnt.tolist(), ends, n_f):
        if c > 0:
            res[n] = (qf[e - m:e], tf[e - m:e], m)
    return res


def _similarities(cfg, tasks, P, items):
    """matching.py:194-202 for items [(task index, (q, t, n_in))]: similarity RANSAC at
    0.02 max(W_i, H_i), then the least-squares fit on its inliers (both on the device); None
    when it fails."""
    res = [None] * len(items)
    ok = [n for n, (_, (q, _, _)) in enumerate(items) if len(q) >= 2]
    _, _, cnt, fits = ransac_rows("similarity", P, [P.rows(tasks[items[n][0]][0], items[n][1][0]) for n in ok],
                                  [P.rows(tasks[items[n][0]][1], items[n][1][1]) for n in ok],
                                  [0.02 * cfg.side(tasks[items[n][0]][0]) for n in ok],
                                  [pair_seed(*tasks[items[n][0]], 0x5) for n in ok], S_ITERS, fit=True,
                                  conf=S_CONF)
    good = ((cnt >= 2) & np.isfinite(fits).all(1)).tolist() if ok else []
    for n, g, fit in zip(ok, good, fits.tolist()):
        if g:
            res[n] = tuple(fit)
    return res


# ---- images from feature files --------------------------------------------------------------

def load_images(ctx, idx):
    """FlowImages of photos `idx` from the context's feature files (matching.py:29-64): one
    batched rescale + undistortion launch for all of them.  Keypoint size / angle are read when
    the file holds them (the learned filter's inputs)."""
    torch = _torch()
    from . import camera
    items, meta = [], []
    for i in idx:
        with np.load(os.path.join(ctx["feat_dir"], "%d.npz" % i)) as d:
            xy = np.ascontiguousarray(d["xy"], np.float32).reshape(-1, 2)
            desc = np.ascontiguousarray(d["desc"], np.uint8).reshape(-1, 128)
            rw, rh = (int(v) for v in d["red_wh"])
            size = d["size"].astype(np.float32) if "size" in d.files else None
            angle = d["angle"].astype(np.float32) if "angle" in d.files else None
        W, H = ctx["W"][i], ctx["H"][i]
        K, dd = ctx["K"][i], ctx["dist"][i]
        k8 = (K[0, 0], K[0, 2], K[1, 2], dd[0], dd[1], dd[2], dd[3], dd[4])
        items.append((torch.from_numpy(xy).cuda(), (W / float(max(rw, 1)), H / float(max(rh, 1))), k8))
        meta.append((torch.from_numpy(desc).cuda(), W, H, size, angle))
    outs = camera.undistort_batch(items, want_xy=False) if items else []
    res = {}
    for i, (_, und), (desc, W, H, size, angle) in zip(idx, outs, meta):
        t = lambda a: torch.from_numpy(a).cuda() if a is not None else None   # noqa: E731
        res[i] = FlowImage(desc, und, W, H, t(size), t(angle))
    return res
