"""Feature extraction front end (replaces the GPU side of app/planesite/features.py).

* ``sift_gpu(g2, nfeat)`` — drop-in of features.sift_gpu (features.py:115-129): detect
  ``int(nfeat*2.2)`` keypoints, grid-spread to ``nfeat``, sort by response, describe.
* ``_spread`` / ``_save`` — the reference's grid spread (features.py:38-57) and RootSIFT npz
  writer (features.py:74-86), same formats.
* ``FeatureBatch`` — the whole flight on the GPU without host round trips: one call detects,
  spreads and describes every image (device-side spread, RootSIFT uint8 epilogue), with the
  pyramid launched once per (octave, layer) for all images of a shape.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib, gsift


class _KP:
    __slots__ = ("pt", "response")

    def __init__(self, x, y, r):
        self.pt = (x, y)
        self.response = r


def _spread(kps, W, H, n, gx=12, gy=9):
    """Keep n keypoints spread over a gx x gy grid, best responses per cell first
    (features.py:38-57).  Ties in response among the fill-up keypoints keep detection order."""
    if len(kps) <= n:
        return list(range(len(kps)))
    resp = np.array([k.response for k in kps], np.float32)
    xy = np.array([k.pt for k in kps], np.float32)
    cx = np.clip((xy[:, 0] / W * gx).astype(int), 0, gx - 1)
    cy = np.clip((xy[:, 1] / H * gy).astype(int), 0, gy - 1)
    cell = cy * gx + cx
    order = np.lexsort((-resp, cell))
    cs = cell[order]
    first = np.searchsorted(cs, cs, side="left")
    rank = np.arange(len(cs)) - first
    cap = int(np.ceil(n / (gx * gy)))
    sel = order[rank < cap]
    if len(sel) < n:
        rest = np.setdiff1d(np.arange(len(kps)), sel)
        rest = rest[np.argsort(-resp[rest], kind="stable")][: n - len(sel)]
        sel = np.concatenate([sel, rest])
    return sel[:n].tolist()


def sift_gpu(g2, nfeat, gpu=True):
    """Same outputs as features.sift_gpu: (xy f32 (n,2), desc f32 (n,128) or None, resp f32 (n,),
    describe time)."""
    h, w = g2.shape[:2]
    kps, state = gsift.detect(g2, gpu=gpu, max_kp=int(nfeat * 2.2))
    n = len(kps["x"])
    if n < 10:
        return np.zeros((0, 2), np.float32), None, np.zeros(0, np.float32), 0.0
    t = time.perf_counter()
    lst = [_KP(float(x), float(y), float(r)) for x, y, r in zip(kps["x"], kps["y"], kps["response"])]
    sel = np.array(_spread(lst, w, h, nfeat), np.int64)
    sel = sel[np.argsort(-kps["response"][sel], kind="stable")]
    desc = gsift.describe(state, kps, sel, budget=4e6)
    xy = np.c_[kps["x"][sel], kps["y"][sel]].astype(np.float32)
    return xy, desc, kps["response"][sel].astype(np.float32), time.perf_counter() - t


def rootsift_u8(desc):
    """features.py:80-82 (host numpy): L1-normalise, sqrt, x512, clip, truncate to uint8."""
    desc = desc / (np.abs(desc).sum(1, keepdims=True) + 1e-7)
    desc = np.sqrt(desc)
    return np.clip(desc * 512.0, 0, 255).astype(np.uint8)


def _save(out_npz, xy, desc, resp, w, h, desc_u8=None):
    """RootSIFT-encode descriptors and write the feature file (features.py:74-86 format).
    ``desc_u8`` (already encoded on the GPU) skips the host encoding."""
    if (desc is None and desc_u8 is None) or len(xy) < 10:
        np.savez(out_npz, xy=np.zeros((0, 2), np.float32), desc=np.zeros((0, 128), np.uint8),
                 resp=np.zeros(0, np.float32), red_wh=np.array([w, h], np.int32))
        return 0
    u8 = desc_u8 if desc_u8 is not None else rootsift_u8(desc)
    np.savez(out_npz, xy=np.asarray(xy, np.float32), desc=u8, resp=np.asarray(resp, np.float32),
             red_wh=np.array([w, h], np.int32))
    return len(xy)


class FeatureBatch:
    """Device-resident SIFT features for a list of uint8 images (torch tensors on the GPU).

    ``run()`` enqueues detect -> spread -> describe (+RootSIFT uint8) for every image in one
    C-ABI call; outputs stay on the device until ``results()``.  Images of equal shape share
    their pyramid launches."""

    def __init__(self, shapes, nfeat=8000, max_kp=None, contrast=0.04, edge=10.0, device="cuda",
                 float_desc=False, caps=None):
        import torch
        self.torch = torch
        self.shapes = [tuple(int(v) for v in s) for s in shapes]
        self.n = len(self.shapes)
        self.nfeat = int(nfeat)
        self.max_kp = int(max_kp if max_kp is not None else int(nfeat * 2.2))
        self.contrast = float(contrast)
        self.edge = float(edge)
        Hm = max(s[0] for s in self.shapes)
        Wm = max(s[1] for s in self.shapes)
        self.cand_cap, self.kp_cap = caps if caps else gsift.default_caps(Hm, Wm, self.max_kp)
        self.device = device
        self.float_desc = float_desc
        self._ran = False
        self._prepared = None
        self._graph = None
        self._alloc()

    def _alloc(self):
        torch = self.torch
        dev = self.device
        self.ws, self.ws_bytes = [], []
        for (H, W) in self.shapes:
            nb = C.c_size_t()
            _lib.call("fsx_sift_workspace_bytes", H, W, self.cand_cap, self.kp_cap, C.byref(nb))
            self.ws.append(torch.empty(nb.value, dtype=torch.uint8, device=dev))
            self.ws_bytes.append(nb.value)
        self.kp = torch.empty((self.n, 7, self.max_kp), dtype=torch.float32, device=dev)
        self.n_out = torch.zeros(self.n, dtype=torch.int32, device=dev)
        self.u8 = torch.empty((self.n, max(self.nfeat, 1), 128), dtype=torch.uint8, device=dev)
        self.desc = (torch.empty((self.n, max(self.nfeat, 1), 128), dtype=torch.float32, device=dev)
                     if self.float_desc else None)
        tb = C.c_size_t()
        _lib.call("fsx_sift_table_bytes", self.n, C.byref(tb))
        self.table = torch.empty(tb.value, dtype=torch.uint8, device=dev)
        self.table_bytes = tb.value
        n = self.n
        P = C.c_void_p
        self._c_ws = (P * n)(*[t.data_ptr() for t in self.ws])
        self._c_wsb = (C.c_size_t * n)(*self.ws_bytes)
        self._c_kp = (P * n)(*[self.kp[i].data_ptr() for i in range(n)])
        self._c_u8 = (P * n)(*[self.u8[i].data_ptr() for i in range(n)])
        self._c_desc = (P * n)(*[self.desc[i].data_ptr() for i in range(n)]) if self.desc is not None else None
        self._c_H = (C.c_int * n)(*[s[0] for s in self.shapes])
        self._c_W = (C.c_int * n)(*[s[1] for s in self.shapes])

    STAGES = ("pyramid", "extrema", "refine", "orientation", "select", "spread", "describe")

    def _args(self, c_img, c_ld):
        P = C.c_void_p
        return (self.n, c_img, self._c_H, self._c_W, c_ld, self.contrast, self.edge, self.max_kp, self.nfeat,
                self.cand_cap, self.kp_cap, self._c_ws, self._c_wsb, self._c_kp, P(self.n_out.data_ptr()),
                self._c_desc, self._c_u8, P(self.table.data_ptr()), self.table_bytes)

    def run(self, imgs, events=None, stream=None, graph=False):
        """Enqueue the feature pipeline for device images (list of uint8 (H, W) tensors).
        ``events``: 8 torch.cuda.Event(enable_timing=True) recorded around the stages.
        The job table is uploaded when the image buffers change; otherwise a run only enqueues
        kernels.  ``graph=True`` replays a HIP graph of the launch sequence (captured on first
        use for these buffers; events are then not recorded): refill the same image tensors
        in place to process new photos."""
        n = self.n
        if len(imgs) != n:
            raise ValueError("expected %d images" % n)
        for t, s in zip(imgs, self.shapes):
            if tuple(t.shape) != s or t.dtype != self.torch.uint8 or not t.is_cuda:
                raise ValueError("image shape/dtype/device mismatch")
        P = C.c_void_p
        key = tuple((t.data_ptr(), t.stride(0)) for t in imgs)
        c_img = (P * n)(*[t.data_ptr() for t in imgs])
        c_ld = (C.c_int * n)(*[t.stride(0) for t in imgs])
        st = _lib.stream_handle(stream)
        if key != self._prepared:
            _lib.call("fsx_sift_features_prepare", *self._args(c_img, c_ld), P(st))
            self._prepared = key
            self._graph = None
        self._ran = True
        if graph:
            if self._graph is None:
                self._graph = self._capture(c_img, c_ld)
            if stream is not None:
                with self.torch.cuda.stream(stream):
                    self._graph.replay()
            else:
                self._graph.replay()
            return
        ev = None
        if events is not None:
            for e in events:
                if not e.cuda_event:     # torch creates the HIP event on first record
                    e.record()
            ev = (P * 8)(*[e.cuda_event for e in events])
        _lib.call("fsx_sift_features_launch", *self._args(c_img, c_ld), ev, P(st))

    def _capture(self, c_img, c_ld):
        torch = self.torch
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g, stream=side):
            _lib.call("fsx_sift_features_launch", *self._args(c_img, c_ld), None, C.c_void_p(_lib.stream_handle()))
        torch.cuda.current_stream().wait_stream(side)
        return g

    @classmethod
    def stage_times(cls, events):
        """Milliseconds per stage from the events of one run() (synchronises on the last)."""
        events[-1].synchronize()
        return {k: events[i].elapsed_time(events[i + 1]) for i, k in enumerate(cls.STAGES)}

    def check_capacity(self, stream=None):
        """Raise if any image overflowed a workspace capacity (synchronises)."""
        if not self._ran:
            raise RuntimeError("FeatureBatch.check_capacity() before any run()")
        st = _lib.stream_handle(stream)
        for i, ws in enumerate(self.ws):
            cnt = (C.c_int * 4)()
            _lib.call("fsx_sift_counters", C.c_void_p(ws.data_ptr()), cnt, C.c_void_p(st))
            if cnt[3]:
                raise RuntimeError("image %d overflowed capacity (flags %d, counts %d/%d/%d)" %
                                   (i, cnt[3], cnt[0], cnt[1], cnt[2]))

    def counts(self):
        """(detected, selected) keypoint counts per image (synchronises)."""
        n = self.n_out.cpu().numpy().astype(np.int64)
        return n, np.minimum(n, self.nfeat)

    def results(self, i):
        """Host arrays of image i in features.sift_gpu's output order: xy (m,2) f32,
        RootSIFT desc u8 (m,128), resp (m,), plus the float descriptors if kept."""
        n_det, m = self.counts()
        m = int(m[i])
        kp = self.kp[i].cpu().numpy()
        sel = self._sel(i, m)
        xy = np.c_[kp[0][sel], kp[1][sel]].astype(np.float32)
        resp = kp[4][sel].astype(np.float32)
        u8 = self.u8[i, :m].cpu().numpy()
        d = self.desc[i, :m].cpu().numpy() if self.desc is not None else None
        return xy, u8, resp, d

    def selection_device(self, i):
        """The spread selection of image i as a device int64 tensor of nfeat entries (entries at
        and beyond the selected count are stale)."""
        off = C.c_size_t()
        H, W = self.shapes[i]
        _lib.call("fsx_sift_selection_offset", H, W, self.cand_cap, self.kp_cap, C.byref(off))
        o = off.value
        return self.ws[i][o:o + 8 * max(self.nfeat, 1)].view(self.torch.int64)

    def _sel(self, i, m):
        """The spread selection (int64 indices into the detect arrays, final order) of image i."""
        off = C.c_size_t()
        H, W = self.shapes[i]
        _lib.call("fsx_sift_selection_offset", H, W, self.cand_cap, self.kp_cap, C.byref(off))
        o = off.value
        return self.ws[i][o:o + 8 * m].cpu().numpy().view(np.int64)


# ---- GPU feature worker (features.py:270-381) ---------------------------------------------------

def read_gray(path):
    """Decode one photo for the worker: ``.npy`` files holding a 2-D uint8 gray image (already
    reduced) are read directly.  JPEG decoding, reduction and CLAHE (features.py:60-71) are
    outside this library; other paths raise, and the worker reports that photo as failed."""
    p = str(path)
    if not p.endswith(".npy"):
        raise ValueError("no decoder for %s (the worker reads .npy gray images)" % p.rsplit("/", 1)[-1])
    g = np.load(p, allow_pickle=False)
    if g.dtype != np.uint8 or g.ndim != 2:
        raise ValueError("%s: expected a 2-D uint8 array, got %s %s" % (p.rsplit("/", 1)[-1], g.dtype, g.shape))
    return np.ascontiguousarray(g)


def _gpu_ready():
    """Device name, memory and a short self-check of the library (features.py:179-210's role)."""
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no ROCm GPU visible")
    _lib.load()
    prop = torch.cuda.get_device_properties(0)
    g = (np.add.outer(np.arange(64), np.arange(80)) % 7 * 30).astype(np.uint8)
    gsift.detect(g, max_kp=100)
    return "%s (%.1f GB)" % (prop.name, prop.total_memory / 2 ** 30)


_CLAHE = []


def _clahe():
    """The worker's CLAHE object (clip 2.0, 8 x 8 tiles, features.py:70), created on first use."""
    if not _CLAHE:
        from .clahe import createCLAHE
        _CLAHE.append(createCLAHE(2.0, (8, 8)))
    return _CLAHE[0]


def gpu_worker_main(task_q, res_q, n_threads, wid, batch=16, decode=read_gray, clahe=False):
    """features.py:270-381 with the one-photo GPU loop replaced by batches.  ``n_threads``
    decoder threads take tasks (idx, path, out_npz, thumb_path, reduce, nfeat) from task_q
    (None ends a thread) and decode them; the GPU loop gathers what is ready (at least one,
    at most ``batch`` photos), runs them through one FeatureBatch per nfeat (the reference's
    sift_gpu settings: max_kp = 2.2 nfeat, contrast 0.015, edge 12), writes each feature file
    with _save and posts ("ok", (idx, n, err, timings)).  Also posted: ("gpu", wid, info) when
    ready, ("gpu_off", wid, msg) if the GPU cannot be used (there is no CPU path here: every
    photo is then reported failed), and ("exit", "gpu%d" % wid) at the end.  thumb_path and
    reduce belong to the decoder the reference uses and are not interpreted.  With
    ``clahe=True`` the decoder returns the plain gray image and the contrast equalisation of
    features.py:70-71 (clip 2.0, 8 x 8 tiles) runs on the GPU on each photo of the batch."""
    import queue
    import threading
    try:
        t0 = time.perf_counter()
        info = _gpu_ready()
        res_q.put(("gpu", wid, "%s, ready in %.1f s" % (info, time.perf_counter() - t0)))
        gpu_err = None
    except Exception as e:  # noqa: BLE001 - reported in the protocol, as features.py:299
        gpu_err = "%s: %s" % (type(e).__name__, str(e)[:300])
        res_q.put(("gpu_off", wid, gpu_err))

    ready = queue.Queue(maxsize=max(4, n_threads, batch))

    def decoder():
        while True:
            task = task_q.get()
            if task is None:
                ready.put(None)
                break
            tt = {"read": 0.0, "decode": 0.0, "bytes": 0, "photos_gpu": 1}
            try:
                t1 = time.perf_counter()
                g2 = decode(task[1])
                tt["decode"] = time.perf_counter() - t1
                tt["bytes"] = int(g2.nbytes)
                ready.put((task, g2, tt, None))
            except Exception as e:  # noqa: BLE001 - per-photo error, as features.py:335-336
                ready.put((task, None, tt, "%s: %s" % (type(e).__name__, e)))

    ths = [threading.Thread(target=decoder, daemon=True) for _ in range(n_threads)]
    for t in ths:
        t.start()
    cache = {}

    def run_group(group):
        import torch
        nfeat = int(group[0][0][5])
        shapes = tuple(g2.shape for _, g2, _, _ in group)
        key = (shapes, nfeat)
        fb = cache.pop(key, None)
        if fb is None:
            fb = FeatureBatch(list(shapes), nfeat=nfeat, contrast=0.015, edge=12.0)
        cache[key] = fb
        while len(cache) > 4:
            cache.pop(next(iter(cache)))
        t2 = time.perf_counter()
        dev = [torch.from_numpy(g2).cuda() for _, g2, _, _ in group]
        if clahe:
            dev = [_clahe().apply(d) for d in dev]
        fb.run(dev)
        fb.check_capacity()
        for k, (task, g2, tt, _) in enumerate(group):
            xy, u8, resp, _ = fb.results(k)
            n = _save(task[2], xy, None, resp, g2.shape[1], g2.shape[0], desc_u8=u8)
            tt["gpu_detect"] = (time.perf_counter() - t2) / len(group)
            tt["gpu_describe"] = 0.0
            res_q.put(("ok", (task[0], n, None, tt)))

    def process(items):
        good = []
        for task, g2, tt, err in items:
            if err or gpu_err:
                res_q.put(("ok", (task[0], 0, err or ("GPU unavailable: " + gpu_err), tt)))
            else:
                good.append((task, g2, tt, None))
        by_nfeat = {}
        for it in good:
            by_nfeat.setdefault(int(it[0][5]), []).append(it)
        for group in by_nfeat.values():
            try:
                run_group(group)
            except Exception as e:  # noqa: BLE001 - the batch's photos are reported failed
                msg = "%s: %s" % (type(e).__name__, e)
                for task, _, tt, _ in group:
                    res_q.put(("ok", (task[0], 0, msg, tt)))

    finished = 0
    while finished < n_threads:
        items = []
        item = ready.get()
        while True:
            if item is None:
                finished += 1
            else:
                items.append(item)
            if len(items) >= batch or finished >= n_threads:
                break
            try:
                item = ready.get_nowait()
            except queue.Empty:
                break
        if items:
            process(items)
    for t in ths:
        t.join()
    res_q.put(("exit", "gpu%d" % wid))
