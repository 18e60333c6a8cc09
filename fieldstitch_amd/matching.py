"""Descriptor matcher on the GPU (replaces the matcher core of app/planesite/matching.py).

Drop-ins with the reference's signatures and return types:

* ``_ratio_match_bf(da, db, ratio, na=None, nb=None) -> (q int64, t int64)`` (matching.py:86-100)
* ``_guided(A, B, Hm, radius, kc, ratio) -> (q int64, t int64)`` (matching.py:113-152); A and B
  are the feature dicts of matching._get (``und`` float64 (n,2), ``desc`` (n,128) holding
  uint8 values, ``wh``)
* ``_unique(q, t, dd)`` (matching.py:103-110)
* ``init_worker``, ``load_features``, ``load_xy``, ``_get`` (matching.py:17-64; rescale and
  undistortion on the GPU) and ``_predict(u, i, j)`` (matching.py:155-172, ground-plane prior
  on the GPU)

``PairMatcher`` runs many pairs (the pipeline's match_pair pool, pipeline.py:535-548) in one
batched launch sequence on device-resident descriptors.  Two forms of the contraction:
``form="plain"`` (FP32 gather + VALU dot products, the baseline) and ``form="mfma"``
(spatially blocked INT8 MFMA on the bytes shifted by -128, with the prior-window mask and
top-2 fused).  Both return the
FP64-oracle match set for uint8 RootSIFT descriptors (all distances are exact integers).
"""
from __future__ import annotations

import ctypes as C
import os
from collections import OrderedDict

import numpy as np

from . import _lib

FORMS = {"plain": 0, "mfma": 1}

# ---- worker context (matching.py:13-26 _CTX / init_worker) ----------------------------------
_CTX = None
_CACHE = OrderedDict()
_CACHE_MAX = 48
_FLOW_CACHE = OrderedDict()       # photo -> match_flow.FlowImage (device-resident features)


def init_worker(ctx):
    """matching.py:17-26: install the matching context (the dict the pipeline pickles: feat_dir,
    W, H, K, dist, R, C, zg, ...).  ``ctx`` is that dict or the path of its pickle file."""
    global _CTX
    if isinstance(ctx, (str, os.PathLike)):
        import pickle
        with open(ctx, "rb") as f:
            ctx = pickle.load(f)
    _CTX = ctx
    _CACHE.clear()
    _FLOW_CACHE.clear()


def _read_npz(feat_dir, i):
    d = np.load(os.path.join(feat_dir, "%d.npz" % i))
    return d["xy"].astype(np.float32), d["desc"], d["red_wh"]


def load_features(feat_dir, i, W, H):
    """matching.py:29-35: feature file i -> (xy float64 full-resolution pixels, desc uint8).
    The rescale (xy + 0.5) * (W/rw, H/rh) - 0.5 runs on the GPU (bit-exact float64)."""
    xy32, desc, red = _read_npz(feat_dir, i)
    if len(xy32) == 0:
        return xy32.astype(np.float64), desc
    xy, _ = _rescale_undistort(xy32, red, W, H, (1.0, 0.0, 0.0, 0, 0, 0, 0, 0))
    return xy, desc


def load_xy(feat_dir, i, W, H):
    """matching.py:38-45: positions only."""
    return load_features(feat_dir, i, W, H)[0]


def _rescale_undistort(xy32, red_wh, W, H, k):
    import torch
    from . import camera
    rw, rh = int(red_wh[0]), int(red_wh[1])
    d = torch.from_numpy(np.ascontiguousarray(xy32, np.float32)).cuda()
    (xy, und), = camera.undistort_batch([(d, (W / float(rw), H / float(rh)), k)])
    return xy.cpu().numpy(), und.cpu().numpy()


def _get(i):
    """matching.py:48-64: cached feature dict of image i: xy, und (float64), desc (float32 of the
    uint8 values), norm, tree (None), wh.  Rescale and undistortion run on the GPU."""
    if i in _CACHE:
        _CACHE.move_to_end(i)
        return _CACHE[i]
    c = _CTX
    xy32, desc, red = _read_npz(c["feat_dir"], i)
    if len(xy32):
        K, dd = c["K"][i], c["dist"][i]
        k = (K[0, 0], K[0, 2], K[1, 2], dd[0], dd[1], dd[2], dd[3], dd[4])
        xy, und = _rescale_undistort(xy32, red, c["W"][i], c["H"][i], k)
    else:
        xy = und = xy32.astype(np.float64)
    d = desc.astype(np.float32)
    item = {"xy": xy, "und": und, "desc": d, "norm": (d * d).sum(1), "tree": None, "wh": (c["W"][i], c["H"][i])}
    _CACHE[i] = item
    while len(_CACHE) > _CACHE_MAX:
        _CACHE.popitem(last=False)
    return item


def ground_predict(u, Kinv_i, R_i, C_i, zg_i, R_j, K_j, C_j, stream=None):
    """Ground-plane prior on the GPU: undistorted pixels u (n,2) of camera i -> (pred (n,2)
    float64 pixels in camera j, ok bool) (matching.py:155-172)."""
    import torch
    u = np.ascontiguousarray(np.asarray(u, np.float64).reshape(-1, 2))
    n = len(u)
    if n == 0:
        return np.zeros((0, 2), np.float64), np.zeros(0, bool)
    cams = np.concatenate([np.asarray(Kinv_i, np.float64).ravel(), np.asarray(R_i, np.float64).ravel(),
                           np.asarray(C_i, np.float64).ravel(), [float(zg_i)], np.asarray(R_j, np.float64).ravel(),
                           np.asarray(K_j, np.float64).ravel(), np.asarray(C_j, np.float64).ravel()])
    if cams.size != 43:
        raise ValueError("camera parameters must be 3x3 / 3-vectors")
    d_u = torch.from_numpy(u).cuda()
    pred = torch.empty((n, 2), dtype=torch.float64, device=d_u.device)
    ok = torch.empty(n, dtype=torch.uint8, device=d_u.device)
    c_cams = (C.c_double * 43)(*cams.tolist())
    _lib.call("fsx_ground_predict", n, C.c_void_p(d_u.data_ptr()), c_cams, C.c_void_p(pred.data_ptr()),
              C.c_void_p(ok.data_ptr()), C.c_void_p(_lib.stream_handle(stream)))
    return pred.cpu().numpy(), ok.cpu().numpy().astype(bool)


def _predict(u, i, j):
    """matching.py:155-172: where undistorted pixels u of image i land in image j (ground plane
    prior from the worker context)."""
    c = _CTX
    return ground_predict(u, np.linalg.inv(c["K"][i]), c["R"][i], c["C"][i], c["zg"][i], c["R"][j], c["K"][j],
                          c["C"][j])


def _unique(q, t, dd):
    """matching.py:103-110 (host): one query per target.  Visiting the matches by distance
    (a stable order, so equal distances keep their input position), the first visit of each
    target wins; np.unique returns those first visits by ascending target."""
    if len(q) == 0:
        return q, t
    by_dist = np.argsort(np.asarray(dd), kind="stable")
    _, first = np.unique(np.asarray(t)[by_dist], return_index=True)
    win = by_dist[first]
    return q[win], t[win]


def as_u8(desc) -> np.ndarray:
    """Descriptors as uint8 (the pipeline's RootSIFT format); float input must hold exact
    integers in 0..255 (matching._get's desc.astype(float32))."""
    d = np.asarray(desc)
    if d.dtype == np.uint8:
        return np.ascontiguousarray(d)
    u = np.clip(d, 0, 255).astype(np.uint8)
    if not np.array_equal(u.astype(d.dtype), d):
        raise ValueError("descriptors must be uint8-valued (RootSIFT u8); float descriptors are not supported")
    return np.ascontiguousarray(u)


class PairMatcher:
    """Batched guided / brute-force matching of image pairs on one GPU.

    images: list of (desc uint8 (n,128), xy float64 (n,2), (W, H)) — numpy or torch tensors.
    pairs:  list of (i, j); Hm: list of 3x3 prior mappings i px -> j px (None = brute force).
    """

    def __init__(self, images, pairs, Hm=None, radius=60.0, kc=0, ratio=0.8, device="cuda"):
        import torch
        self.torch = torch
        self.n_img = len(images)
        self.pairs = [(int(a), int(b)) for a, b in pairs]
        self.n_pair = len(self.pairs)
        self.radius = float(radius)
        self.kc = int(kc)
        self.ratio = float(ratio)
        descs, xys, self.nk, self.W, self.H = [], [], [], [], []
        for d, xy, (W, H) in images:
            d = d if isinstance(d, torch.Tensor) else torch.from_numpy(as_u8(d))
            xy = xy if isinstance(xy, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(xy, np.float64))
            if d.dtype != torch.uint8 or xy.dtype != torch.float64:
                raise TypeError("descriptors must be uint8 and positions float64")
            descs.append(d.reshape(-1, 128).to(device))
            xys.append(xy.reshape(-1, 2).to(device))
            self.nk.append(int(d.shape[0]))
            self.W.append(int(W))
            self.H.append(int(H))
        self.desc = torch.cat(descs).contiguous() if descs else torch.empty((0, 128), dtype=torch.uint8, device=device)
        self.xy = torch.cat(xys).contiguous() if xys else torch.empty((0, 2), dtype=torch.float64, device=device)
        if self.desc.shape[0] == 0:      # keep a valid device pointer
            self.desc = torch.zeros((1, 128), dtype=torch.uint8, device=device)
            self.xy = torch.zeros((1, 2), dtype=torch.float64, device=device)
        n = self.n_img
        self._nk = (C.c_int * n)(*self.nk)
        self._W = (C.c_int * n)(*self.W)
        self._H = (C.c_int * n)(*self.H)
        npair = max(self.n_pair, 1)
        self._pa = (C.c_int * npair)(*[a for a, _ in self.pairs])
        self._pb = (C.c_int * npair)(*[b for _, b in self.pairs])
        dense = [1 if (Hm is None or Hm[p] is None) else 0 for p in range(self.n_pair)]
        self._dense = (C.c_int * npair)(*dense)
        hm = np.zeros((npair, 9), np.float64)
        for p in range(self.n_pair):
            if not dense[p]:
                hm[p] = np.asarray(Hm[p], np.float64).reshape(9)
        self._hm = np.ascontiguousarray(hm)
        nb = C.c_size_t()
        cap = C.c_longlong()
        _lib.call("fsx_match_workspace_bytes", n, self._nk, self._W, self._H, self.n_pair, self._pa, self._pb,
                  self.radius, C.byref(nb), C.byref(cap))
        self.ws = torch.empty(nb.value, dtype=torch.uint8, device=device)
        self.ws_bytes = nb.value
        self.capacity = cap.value
        self._offsets = (C.c_longlong * npair)()
        self._info = (C.c_longlong * 64)()
        self._graphs = {}
        P = C.c_void_p
        # plan + table upload once; every run() only enqueues kernels
        _lib.call("fsx_match_prepare", self.n_img, self._nk, self._W, self._H, P(self.desc.data_ptr()),
                  P(self.xy.data_ptr()), self.n_pair, self._pa, self._pb, self._dense,
                  self._hm.ctypes.data_as(C.POINTER(C.c_double)), self.radius, self.kc, self.ratio,
                  P(self.ws.data_ptr()), self.ws_bytes, self._offsets, self._info,
                  P(_lib.stream_handle()))

    def run(self, form="mfma", events=None, stream=None, graph=False):
        """Enqueue matching of every pair (asynchronous).  ``graph=True`` replays a HIP graph of
        the launch sequence (captured on first use per form); events are then not recorded."""
        if graph:
            g = self._graphs.get(form) or self.capture(form)
            if stream is not None:
                with self.torch.cuda.stream(stream):
                    g.replay()
            else:
                g.replay()
            return
        P = C.c_void_p
        st = _lib.stream_handle(stream)
        ev = None
        if events is not None:
            for e in events:
                if not e.cuda_event:
                    e.record()
            ev = (P * 4)(*[e.cuda_event for e in events])
        _lib.call("fsx_match_launch", P(self.ws.data_ptr()), self._info, FORMS[form], ev, P(st))

    def run_filtered(self, form, net, tau, sizes, angles, stream=None):
        """run() with the learned match filter as a pruning stage (asynchronous up to one count
        read): contraction without publication, then the filter scores every query that passed
        the ratio test (its row built from the query and its best candidate), queries scoring
        below tau are dropped, and the one-query-per-target rule and compaction run on the
        rest.  net: filter_mlp.MatchFilterMLP; sizes / angles: per image float32 keypoint size
        and angle (degrees), as filter_features takes them.  net = None keeps every query (the
        result is run()'s)."""
        torch = self.torch
        P = C.c_void_p
        st = P(_lib.stream_handle(stream))
        sz, an = self._size_angle(sizes, angles) if net is not None else (None, None)
        nq = sum(self.nk[a] for a, _ in self.pairs)
        lst = torch.empty((max(nq, 1), 2), dtype=torch.int32, device=self.desc.device)
        cnt = torch.empty(1, dtype=torch.int32, device=self.desc.device)
        _lib.call("fsx_match_contract", P(self.ws.data_ptr()), self._info, FORMS[form], st)
        _lib.call("fsx_match_accepted", P(self.ws.data_ptr()), self._info, P(lst.data_ptr()), P(cnt.data_ptr()), st)
        n_acc = int(cnt.item())
        score = None
        if n_acc and net is not None:
            X = torch.empty((n_acc, 288), dtype=torch.bfloat16, device=self.desc.device)
            _lib.call("fsx_match_filter_rows", P(self.ws.data_ptr()), self._info, P(sz.data_ptr()), P(an.data_ptr()),
                      P(lst.data_ptr()), P(cnt.data_ptr()), n_acc, P(X.data_ptr()), st)
            score = net.predict(X)
        self._filter_kept = None if score is None else int((score >= tau).sum())
        self._filter_scored = n_acc
        _lib.call("fsx_match_publish", P(self.ws.data_ptr()), self._info, P(lst.data_ptr()), P(cnt.data_ptr()),
                  n_acc, P(score.data_ptr()) if score is not None else None, float(tau), st)

    def filter_counts(self):
        """(queries scored, queries kept) of the last run_filtered."""
        return self._filter_scored, self._filter_kept

    def _size_angle(self, sizes, angles):
        torch = self.torch

        def cat(vals):
            if any(v is None for v in vals):
                raise ValueError("the learned filter needs every image's keypoint sizes and angles")
            return torch.cat([v.float() if isinstance(v, torch.Tensor) else torch.as_tensor(np.asarray(v, np.float32))
                              for v in vals]).to(self.desc.device).contiguous()
        sz, an = cat(sizes), cat(angles)
        if sz.numel() != sum(self.nk) or an.numel() != sum(self.nk):
            raise ValueError("sizes / angles must give one value per keypoint")
        return sz, an

    def candidates(self):
        """Device views of the last run's per-query arrays (caller's query order, pair p's
        queries at the prefix of n_a): best candidate before the ratio test (int32, -1 none),
        its distance and the second smallest (float32 (n, 2)), accepted target (int32)."""
        torch = self.torch
        offs = (C.c_longlong * 4)()
        _lib.call("fsx_match_candidate_offsets", self.n_img, self._nk, self._W, self._H, self.n_pair, self._pa,
                  self._pb, self.radius, offs)
        nq = sum(self.nk[a] for a, _ in self.pairs)
        cb = self.ws[offs[0]:offs[0] + 4 * nq].view(torch.int32)
        c12 = self.ws[offs[1]:offs[1] + 8 * nq].view(torch.float32).view(nq, 2)
        best = self.ws[offs[2]:offs[2] + 4 * nq].view(torch.int32)
        return cb, c12, best

    def filter_features(self, sizes, angles):
        """The learned filter's input rows (bf16 [n_query, 288] device tensor) for the last run:
        one per query, built from its best candidate before the ratio test (zero rows for
        queries without one).  sizes / angles: per image float32 arrays of keypoint size and
        angle (degrees), in the order of the images given to the constructor."""
        torch = self.torch
        sz, an = self._size_angle(sizes, angles)
        nq = sum(self.nk[a] for a, _ in self.pairs)
        X = torch.empty((max(nq, 1), 288), dtype=torch.bfloat16, device=self.desc.device)
        _lib.call("fsx_match_filter_features", C.c_void_p(self.ws.data_ptr()), self._info, C.c_void_p(sz.data_ptr()),
                  C.c_void_p(an.data_ptr()), C.c_void_p(X.data_ptr()), C.c_void_p(_lib.stream_handle()))
        return X[:nq]

    def query_offsets(self):
        """First global query index of each pair (the row offsets of filter_features)."""
        out, o = [], 0
        for a, _ in self.pairs:
            out.append(o)
            o += self.nk[a]
        return out

    def capture(self, form="mfma"):
        """Capture the launch sequence of `form` into a HIP graph (torch.cuda.CUDAGraph)."""
        torch = self.torch
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            _lib.call("fsx_match_launch", C.c_void_p(self.ws.data_ptr()), self._info, FORMS[form], None,
                      C.c_void_p(_lib.stream_handle()))
        torch.cuda.current_stream().wait_stream(side)
        self._graphs[form] = g
        return g

    def results_dev(self, stream=None):
        """The last run's lists packed on the device: (counts (n_pair,) int64, row offsets
        (n_pair + 1,) int64, q, t int32 device tensors) with pair p's (q, t) in rows off[p] ..
        off[p + 1], sorted by t.  Synchronises once, for the counts; only the matches are
        moved (fsx_match_pack), not the per-query output slots."""
        import torch
        ts = torch.cuda.current_stream() if stream is None else stream
        st = _lib.stream_handle(ts)
        npair = max(self.n_pair, 1)
        counts = (C.c_int * npair)()
        flags = (C.c_int * 4)()
        _lib.call("fsx_match_results", self.n_img, self._nk, self._W, self._H, self.n_pair, self._pa, self._pb,
                  self.radius, C.c_void_p(self.ws.data_ptr()), counts, None, None, flags, C.c_void_p(st))
        if flags[2]:
            raise RuntimeError("matcher chunk table overflow (flags %d)" % flags[2])
        cnt = np.ctypeslib.as_array(counts)[:self.n_pair].astype(np.int64)
        off = np.zeros(self.n_pair + 1, np.int64)
        np.cumsum(cnt, out=off[1:])
        tot = int(off[-1])
        with torch.cuda.stream(ts):
            q = torch.empty(max(tot, 1), dtype=torch.int32, device=self.ws.device)
            t = torch.empty(max(tot, 1), dtype=torch.int32, device=self.ws.device)
            if tot:
                self._pack_off = torch.from_numpy(off[:-1].copy()).to(self.ws.device)
                _lib.call("fsx_match_pack", C.c_void_p(self.ws.data_ptr()), self._info,
                          C.c_void_p(self._pack_off.data_ptr()), tot, C.c_void_p(q.data_ptr()),
                          C.c_void_p(t.data_ptr()), C.c_void_p(st))
        return cnt, off, q[:tot], t[:tot]

    def results(self, stream=None):
        """List of (q int64, t int64) per pair, sorted by t (synchronises): the lists packed on
        the device, copied to the host once and handed out as views."""
        import torch
        if self.n_pair == 0:
            return []
        _, off, q, t = self.results_dev(stream)
        with torch.cuda.stream(torch.cuda.current_stream() if stream is None else stream):
            qh = q.cpu().numpy().astype(np.int64)
            th = t.cpu().numpy().astype(np.int64)
        o = off.tolist()
        return [(qh[o[p]:o[p + 1]], th[o[p]:o[p + 1]]) for p in range(self.n_pair)]


def _query_state(m: "PairMatcher", stream=None):
    """Per-query (best, b1, cell) arrays of the last run (tests and diagnostics)."""
    nq = sum(m.nk[a] for a, _ in m.pairs)
    best = np.zeros(max(nq, 1), np.int32)
    b1 = np.zeros(max(nq, 1), np.float64)
    cell = np.zeros(max(nq, 1), np.int32)
    _lib.call("fsx_match_query_state", m.n_img, m._nk, m._W, m._H, m.n_pair, m._pa, m._pb, m.radius,
              C.c_void_p(m.ws.data_ptr()), best.ctypes.data_as(C.POINTER(C.c_int)),
              b1.ctypes.data_as(C.POINTER(C.c_double)), cell.ctypes.data_as(C.POINTER(C.c_int)),
              C.c_void_p(_lib.stream_handle(stream)))
    return best[:nq], b1[:nq], cell[:nq]


def _is_u8_valued(d) -> bool:
    d = np.asarray(d)
    if d.dtype == np.uint8:
        return True
    return bool(np.all((d >= 0) & (d <= 255) & (d == np.rint(d))))


def _ratio_match_bf(da, db, ratio, na=None, nb=None, form="mfma"):
    """matching.py:86-100: exact nearest-neighbour ratio matching of all of da against db.
    uint8-valued descriptors (RootSIFT files) take the integer path (``form`` "mfma" or
    "plain"); other float descriptors take FloatMatcher (``form`` "bf16", "fp8", or "mfma" =
    "bf16").  Both return the FP64-oracle match set."""
    if len(da) < 2 or len(db) < 2:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    if form in ("mfma", "plain") and _is_u8_valued(da) and _is_u8_valued(db):
        a, b = as_u8(da), as_u8(db)
        zero = lambda n: np.zeros((n, 2), np.float64)   # noqa: E731
        m = PairMatcher([(a, zero(len(a)), (1, 1)), (b, zero(len(b)), (1, 1))], [(0, 1)], None, radius=1.0,
                        ratio=ratio)
        m.run(form)
        return m.results()[0]
    fm = FloatMatcher([np.asarray(da, np.float32), np.asarray(db, np.float32)], [(0, 1)], ratio=ratio,
                      precision="bf16" if form in ("mfma", "plain") else form)
    fm.run()
    return fm.results()[0]


FLOAT_PRECISION = {"bf16": 1, "fp8": 2}


class FloatMatcher:
    """Batched brute-force ratio matching of float descriptor sets (matching.py:86-110 on float
    descriptors) with MFMA screening in BF16 or FP8 e4m3 (per-descriptor scales), float64
    rescoring of the K best screened targets and a certificate; queries the certificate cannot
    decide are recomputed exactly.  The match set equals the FP64 oracle.

    sets: list of float32 (n, 128) arrays or device tensors; pairs: list of (a, b)."""

    def __init__(self, sets, pairs, ratio=0.8, precision="bf16", K=8, device="cuda"):
        import torch
        self.torch = torch
        if precision not in FLOAT_PRECISION:
            raise ValueError("precision must be 'bf16' or 'fp8'")
        self.precision = precision
        self.K = int(K)
        self.ratio = float(ratio)
        self.pairs = [(int(a), int(b)) for a, b in pairs]
        self.n_pair = len(self.pairs)
        ds = []
        for d in sets:
            d = d if isinstance(d, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(d, np.float32))
            if d.dtype != torch.float32:
                raise TypeError("descriptors must be float32")
            ds.append(d.reshape(-1, 128).to(device))
        self.nk = [int(d.shape[0]) for d in ds]
        self.desc = torch.cat(ds).contiguous() if sum(self.nk) else torch.zeros((1, 128), dtype=torch.float32,
                                                                                 device=device)
        if sum(self.nk) and not bool(torch.isfinite(self.desc).all()):
            raise ValueError("descriptors must be finite")
        n = len(ds)
        npair = max(self.n_pair, 1)
        self._nk = (C.c_int * n)(*self.nk)
        self._pa = (C.c_int * npair)(*[a for a, _ in self.pairs])
        self._pb = (C.c_int * npair)(*[b for _, b in self.pairs])
        self.n_set = n
        nb = C.c_size_t()
        cap = C.c_longlong()
        _lib.call("fsx_fmatch_workspace_bytes", n, self._nk, self.n_pair, self._pa, self._pb, C.byref(nb),
                  C.byref(cap))
        self.ws = torch.empty(nb.value, dtype=torch.uint8, device=device)
        self.capacity = cap.value
        self._offsets = (C.c_longlong * npair)()

    def run(self, events=None, stream=None):
        """Enqueue the batch (asynchronous).  events: 5 timing events (quantize / screen /
        rescore + fallback / unique + compact)."""
        P = C.c_void_p
        ev = None
        if events is not None:
            for e in events:
                if not e.cuda_event:
                    e.record()
            ev = (P * 5)(*[e.cuda_event for e in events])
        _lib.call("fsx_fmatch_batch", self.n_set, self._nk, P(self.desc.data_ptr()), self.n_pair, self._pa,
                  self._pb, self.ratio, FLOAT_PRECISION[self.precision], self.K, P(self.ws.data_ptr()),
                  self.ws.numel(), self._offsets, ev, P(_lib.stream_handle(stream)))

    def results(self, stream=None):
        """List of (q int64, t int64) per pair, sorted by t (synchronises)."""
        npair = max(self.n_pair, 1)
        counts = (C.c_int * npair)()
        q = np.zeros(max(self.capacity, 1), np.int32)
        t = np.zeros(max(self.capacity, 1), np.int32)
        st = (C.c_double * 4)()
        _lib.call("fsx_fmatch_results", self.n_set, self._nk, self.n_pair, self._pa, self._pb,
                  C.c_void_p(self.ws.data_ptr()), counts, q.ctypes.data_as(C.POINTER(C.c_int)),
                  t.ctypes.data_as(C.POINTER(C.c_int)), st, C.c_void_p(_lib.stream_handle(stream)))
        self._stats = {"fallback_queries": int(st[0]), "max_abs_distance_error": float(st[1]),
                       "mean_abs_distance_error": float(st[2]), "rescored_candidates": int(st[3])}
        out = []
        for p in range(self.n_pair):
            o = self._offsets[p]
            out.append((q[o:o + counts[p]].astype(np.int64), t[o:o + counts[p]].astype(np.int64)))
        return out

    def stats(self):
        """Screening error against float64 over the rescored candidates and the number of
        queries decided by the exact fallback (of the last results() call)."""
        return dict(self._stats)


def fp8_layout_probe():
    """Run the 32x32x64 FP8 MFMA layout probe; returns (64 lanes, 16) float32."""
    import torch
    out = torch.zeros(64 * 16, dtype=torch.float32, device="cuda")
    _lib.call("fsx_fp8_layout_probe", C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
    return out.view(64, 16).cpu().numpy()


def fp8_scale_probe():
    """Run the FP8 MFMA probe with per-lane E8M0 scales; returns (64 lanes, 16) float32."""
    import torch
    out = torch.zeros(64 * 16, dtype=torch.float32, device="cuda")
    _lib.call("fsx_fp8_scale_probe", C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
    return out.view(64, 16).cpu().numpy()


def fp8_cvt_probe():
    """The hardware e4m3 conversion the FP8 quantization uses against a software
    round-to-nearest-even conversion on every float in [-448, 448]: (mismatches, smallest
    mismatching bit pattern or None)."""
    import torch
    out = torch.zeros(2, dtype=torch.int64, device="cuda")
    _lib.call("fsx_fp8_cvt_probe", C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
    bad, first = out.cpu().numpy().view(np.uint64).tolist()
    return int(bad), (None if first == 2 ** 64 - 1 else int(first))


def _guided(A, B, Hm, radius, kc, ratio, form="mfma"):
    """matching.py:113-152: match A's features to B's keypoints within `radius` of their
    predicted position Hm(A.und), the kc spatially nearest (kc >= len(B) or 0: all)."""
    nb = len(B["und"])
    kc = 0 if kc >= nb else int(kc)
    m = PairMatcher([(as_u8(A["desc"]), np.asarray(A["und"], np.float64), A["wh"]),
                     (as_u8(B["desc"]), np.asarray(B["und"], np.float64), B["wh"])],
                    [(0, 1)], [np.asarray(Hm, np.float64)], radius=radius, kc=kc, ratio=ratio)
    m.run(form)
    return m.results()[0]


def mfma_i8_layout_probe():
    """Run the 32x32x32 INT8 MFMA layout probe; returns (64 lanes, 16) int32."""
    import torch
    out = torch.zeros(64 * 16, dtype=torch.int32, device="cuda")
    _lib.call("fsx_mfma_i8_layout_probe", C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
    return out.view(64, 16).cpu().numpy()


def mfma_layout_probe():
    """Run the 32x32x16 BF16 MFMA layout probe; returns (64 lanes, 16) float32."""
    import torch
    out = torch.zeros(64 * 16, dtype=torch.float32, device="cuda")
    _lib.call("fsx_mfma_layout_probe", C.c_void_p(out.data_ptr()), C.c_void_p(_lib.stream_handle()))
    return out.view(64, 16).cpu().numpy()


# ---- match_pair (matching.py:175-253) on the batched stage pipeline of match_flow -----------

def _flow_images(idx):
    """FlowImages of photos idx from the worker context's feature files, cached (48 photos, as
    matching.py:13-14)."""
    from . import match_flow
    need = [i for i in dict.fromkeys(idx) if i not in _FLOW_CACHE]
    if need:
        _FLOW_CACHE.update(match_flow.load_images(_CTX, need))
    out = {}
    for i in idx:
        _FLOW_CACHE.move_to_end(i)
        out[i] = _FLOW_CACHE[i]
    while len(_FLOW_CACHE) > max(_CACHE_MAX, len(set(idx))):
        _FLOW_CACHE.popitem(last=False)
    return out


def match_pairs(tasks, form="mfma", filt=None, stats=None):
    """match_pair for a list of pairs of the installed context, each stage batched over all the
    pairs that reach it (fieldstitch_amd.match_flow).  Returns the tuples in task order."""
    from . import match_flow
    tasks = [(int(i), int(j)) for i, j in tasks]
    idx = list(dict.fromkeys(i for t in tasks for i in t))
    bad = {}
    try:
        images = _flow_images(idx)
    except Exception:  # noqa: BLE001 - find the photos that fail to load, report their pairs
        images = {}
        for i in idx:
            try:
                images.update(_flow_images([i]))
            except Exception as e:  # noqa: BLE001
                bad[i] = e
    good = [k for k, (i, j) in enumerate(tasks) if i not in bad and j not in bad]
    res = match_flow.run_pairs([tasks[k] for k in good], images, match_flow.FlowConfig(_CTX), form=form, filt=filt,
                               stats=stats)
    out = [None] * len(tasks)
    for k, r in zip(good, res):
        out[k] = r
    for k, (i, j) in enumerate(tasks):
        if out[k] is None:
            out[k] = match_flow._error_tuple(i, j, bad[i] if i in bad else bad[j])
    return out


def match_pair(task):
    """matching.py:205-253 for the pair task = (i, j) of the installed context: (i, j, q int32,
    t int32, sim, n_inliers), (i, j, None, None, None, n) on failure, or ("error", msg) in the
    sim slot.  One pair through match_pairs; cv2's estimators are replaced by the deterministic
    GPU RANSAC of match_flow, FLANN by the exact brute force."""
    return match_pairs([task])[0]


def _ratio_match(da, db, ratio, checks=48):
    """matching.py:73-83 (FLANN kd-tree, `checks` leaves): the exhaustive GPU ratio matcher
    instead, whose match set is the exact one FLANN approximates."""
    return _ratio_match_bf(da, db, ratio)


def _verify(c, i, j, ui, uj, q, t):
    """matching.py:175-191 for one pair: fundamental-matrix RANSAC at c["ransac_thr"] on the
    GPU -> (q, t of the inliers, count) or (None, None, 0)."""
    from . import match_flow
    cfg = match_flow.FlowConfig(c)
    pos = match_flow.Positions({i: np.asarray(ui, np.float64), j: np.asarray(uj, np.float64)})
    return match_flow._verify(cfg, [(i, j)], pos, [(0, (np.asarray(q), np.asarray(t)))])[0]


def _similarity(c, i, j, ui, uj, q, t):
    """matching.py:194-202 for one pair: (angle rad, scale, tx, ty) of the similarity from
    image i to image j, or None (similarity RANSAC on the GPU, then its least-squares fit)."""
    from . import match_flow
    try:
        cfg = match_flow.FlowConfig(c)
        pos = match_flow.Positions({i: np.asarray(ui, np.float64), j: np.asarray(uj, np.float64)})
        return match_flow._similarities(cfg, [(i, j)], pos, [(0, (np.asarray(q), np.asarray(t), 0))])[0]
    except Exception:  # noqa: BLE001 - the reference returns None on any failure here
        return None
