"""Feature extraction front end (replaces the GPU side of app/planesite/features.py).

* ``sift_gpu(g2, nfeat)`` — drop-in of features.sift_gpu (features.py:115-129) on a one-photo
  FeatureBatch: detect ``int(nfeat*2.2)`` keypoints, grid-spread to ``nfeat`` on the device,
  sort by response, describe.
* ``spread_select`` / ``_spread`` — the grid spread (features.py:38-57) on arrays / keypoint
  objects; ``rootsift_u8`` / ``_save`` — the RootSIFT encoding and the feature file
  (features.py:74-86, same keys, plus optional keypoint size / angle).
* ``FeatureBatch`` — the whole flight on the GPU without host round trips: one call detects,
  spreads and describes every image (device-side spread, RootSIFT uint8 epilogue), with the
  pyramid launched once per (octave, layer) for all images of a shape.
"""
from __future__ import annotations

import ctypes as C
import os
import time
from collections import OrderedDict

import numpy as np

from . import _lib, gsift


def spread_select(x, y, resp, W, H, n, gx=12, gy=9):
    """The grid spread of features.py:38-57 on arrays: indices of at most n keypoints.  Each
    of the gx x gy cells contributes its ceil(n / (gx gy)) strongest keypoints, listed cell by
    cell (row-major cells, strongest first inside a cell); when that is short of n, the
    strongest of the rest follow.  Equal responses keep detection order."""
    m = len(resp)
    if m <= n:
        return np.arange(m, dtype=np.int64)
    resp = np.asarray(resp, np.float32)
    x = np.asarray(x, np.float32)
    y = np.asarray(y, np.float32)
    # cell of each keypoint, in float32 as the reference computes it (x / W * gx, truncated)
    cell = (np.clip((y / H * gy).astype(np.int64), 0, gy - 1) * gx +
            np.clip((x / W * gx).astype(np.int64), 0, gx - 1))
    strongest = np.argsort(-resp, kind="stable")                       # response order, ties by index
    by_cell = strongest[np.argsort(cell[strongest], kind="stable")]    # cell, then response order
    per_cell = np.bincount(cell, minlength=gx * gy)
    first_in_cell = np.cumsum(per_cell) - per_cell
    rank = np.arange(m) - first_in_cell[cell[by_cell]]
    quota = -(-n // (gx * gy))
    chosen = by_cell[rank < quota]
    if len(chosen) < n:
        taken = np.zeros(m, bool)
        taken[chosen] = True
        chosen = np.concatenate([chosen, strongest[~taken[strongest]][:n - len(chosen)]])
    return chosen[:n].astype(np.int64)


def _spread(kps, W, H, n, gx=12, gy=9):
    """features.py:38-57 drop-in: the spread of keypoint objects with .pt and .response
    (cv2.KeyPoint or alike), as a list of indices."""
    pts = np.array([k.pt for k in kps], np.float32).reshape(-1, 2)
    resp = np.array([k.response for k in kps], np.float32)
    return spread_select(pts[:, 0], pts[:, 1], resp, W, H, n, gx, gy).tolist()


_SIFT_BATCHES = OrderedDict()


def _single_batch(shape, nfeat, caps=None):
    """A one-photo FeatureBatch for sift_gpu (gsift.detect's thresholds, float descriptors
    kept), cached per (shape, nfeat, caps)."""
    key = (tuple(shape), int(nfeat), caps)
    fb = _SIFT_BATCHES.pop(key, None)
    if fb is None:
        fb = FeatureBatch([shape], nfeat=nfeat, contrast=0.015, edge=12.0, float_desc=True, caps=caps)
    _SIFT_BATCHES[key] = fb
    while len(_SIFT_BATCHES) > 2:
        _SIFT_BATCHES.popitem(last=False)
    return fb


def sift_gpu(g2, nfeat, gpu=True):
    """features.py:115-129 drop-in on the batched pipeline: one FeatureBatch run detects
    int(2.2 nfeat) keypoints (contrast 0.015, edge 12), spreads them to nfeat on the device,
    orders them by response and describes them.  Returns (xy float32 (n, 2), desc float32
    (n, 128) or None, resp float32 (n,), seconds spent in spread + describe)."""
    if not gpu:
        raise ValueError("fieldstitch_amd.features.sift_gpu runs on the GPU only")
    import torch
    img = g2 if isinstance(g2, torch.Tensor) else torch.from_numpy(np.ascontiguousarray(g2, np.uint8))
    img = img.cuda() if not img.is_cuda else img
    caps = None
    while True:
        fb = _single_batch(tuple(img.shape), nfeat, caps)
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(8)]
        fb.run([img], events=ev)
        try:
            fb.check_capacity()
            break
        except RuntimeError:
            # a capacity overflowed: double both capacities and run again
            caps = tuple(2 * c for c in (caps or (fb.cand_cap, fb.kp_cap)))
    t = FeatureBatch.stage_times(ev)
    n_det, _ = fb.counts()
    if n_det[0] < 10:
        return np.zeros((0, 2), np.float32), None, np.zeros(0, np.float32), 0.0
    xy, _, resp, desc = fb.results(0)
    return xy, desc, resp, (t["spread"] + t["describe"]) / 1e3


def rootsift_u8(desc):
    """The RootSIFT encoding of features.py:80-82 in float32: each row over its L1 norm (+1e-7),
    square root, times 512, clipped to [0, 255], truncated to uint8."""
    d = np.asarray(desc, np.float32)
    l1 = np.add.reduce(np.abs(d), axis=1, keepdims=True) + np.float32(1e-7)
    return np.clip(np.sqrt(d / l1) * np.float32(512.0), 0, 255).astype(np.uint8)


def _save(out_npz, xy, desc, resp, w, h, desc_u8=None, size=None, angle=None):
    """Write one photo's feature file in the format of features.py:74-86: xy float32 (n, 2)
    reduced-image pixels, desc uint8 RootSIFT (n, 128), resp float32 (n,), red_wh int32 [w, h];
    fewer than 10 features (or no descriptors) give empty arrays.  desc_u8: rows already
    encoded on the GPU.  size / angle (optional, float32 (n,)) are stored too for the learned
    match filter; readers of the reference format ignore them.  Returns the feature count."""
    have = (desc is not None or desc_u8 is not None) and len(xy) >= 10
    n = len(xy) if have else 0
    fields = {"xy": np.asarray(xy, np.float32).reshape(-1, 2)[:n],
              "desc": (desc_u8 if desc_u8 is not None else rootsift_u8(desc)) if have else np.zeros((0, 128), np.uint8),
              "resp": np.asarray(resp, np.float32)[:n] if have else np.zeros(0, np.float32),
              "red_wh": np.array([w, h], np.int32)}
    if have and size is not None and angle is not None:
        fields["size"] = np.asarray(size, np.float32)[:n]
        fields["angle"] = np.asarray(angle, np.float32)[:n]
    np.savez(out_npz, **fields)
    return n


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

    def size_angle(self, i):
        """Keypoint size and angle (float32 host arrays) of image i's selected features, in the
        order of results(i)."""
        _, m = self.counts()
        m = int(m[i])
        kp = self.kp[i, 2:4].cpu().numpy()
        sel = self._sel(i, m)
        return kp[0][sel].astype(np.float32), kp[1][sel].astype(np.float32)

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

def read_gray(path, reduce=1):
    """Decode one photo to a 2-D uint8 gray image (features.py:18-26, decode(buf, reduce,
    color=False)).  ``.npy``: a stored gray image, already reduced (``reduce`` is not applied).
    Anything else goes through Pillow: a JPEG is decoded reduced by 1, 2, 4 or 8 in the DCT
    domain (libjpeg's scaled decode, which cv2's IMREAD_REDUCED_GRAYSCALE_<r> uses for JPEG;
    the output is ceil(w / r) x ceil(h / r)) to its luma plane, with EXIF orientation ignored
    as IMREAD_IGNORE_ORIENTATION asks; other formats Pillow reads are decoded at full size and
    only with reduce 1.  Missing or unreadable files raise, and the worker reports that photo as
    failed (features.py:335-336)."""
    p = str(path)
    if p.endswith(".npy"):
        g = np.load(p, allow_pickle=False)
        if g.dtype != np.uint8 or g.ndim != 2:
            raise ValueError("%s: expected a 2-D uint8 array, got %s %s" % (p.rsplit("/", 1)[-1], g.dtype, g.shape))
        return np.ascontiguousarray(g)
    return _pil_decode(p, reduce, "L")


def _pil_decode(path, reduce, mode):
    """Pillow decode of `path` to `mode` ("L" or "RGB"), JPEG reduced in the DCT domain."""
    from PIL import Image
    if reduce not in (1, 2, 4, 8):
        raise ValueError("reduce must be 1, 2, 4 or 8 (got %r)" % (reduce,))
    with Image.open(path) as im:
        w, h = im.size
        if im.format == "JPEG":
            if reduce > 1 or im.mode != mode:
                # draft takes the largest libjpeg scale 1/s with s <= w // (requested width)
                # (and height alike): asking for the floor sizes gives s = reduce, and libjpeg's
                # output is ceil(w / reduce) x ceil(h / reduce)
                im.draft(mode, (max(1, w // reduce), max(1, h // reduce)))
        elif reduce != 1:
            raise ValueError("%s: reduced decoding is for JPEG photos (%s given)" % (path.rsplit("/", 1)[-1], im.format))
        out = np.asarray(im.convert(mode), dtype=np.uint8)
    exp = ((h + reduce - 1) // reduce, (w + reduce - 1) // reduce)
    if out.shape[:2] != exp:
        raise ValueError("%s: decoded %s, expected %s" % (path.rsplit("/", 1)[-1], out.shape[:2], exp))
    return np.ascontiguousarray(out)


def prepare_task(task):
    """The reference's prepare (features.py:60-71) for a worker task (idx, path, out_npz,
    thumb_path, reduce, nfeat), without its CLAHE step (run the worker with ``clahe=True`` so the
    GPU equalises each batch): if ``thumb_path`` is given and missing, the photo is decoded in
    colour, a thumbnail of ((w + 1) * reduce // 8) x ((h + 1) * reduce // 8) (area average;
    the colour image itself at reduce 8) is written as JPEG quality 88 and the gray image is its
    luma; otherwise the gray image is decoded directly (read_gray)."""
    path, thumb, reduce = str(task[1]), task[3], int(task[4])
    if thumb and not os.path.exists(thumb) and not path.endswith(".npy"):
        from PIL import Image
        col = _pil_decode(path, reduce, "RGB")
        im = Image.fromarray(col)
        if reduce < 8:
            im = im.resize(((col.shape[1] + 1) * reduce // 8, (col.shape[0] + 1) * reduce // 8), Image.BOX)
        im.save(thumb, format="JPEG", quality=88)
        return np.ascontiguousarray(np.asarray(Image.fromarray(col).convert("L"), dtype=np.uint8))
    return read_gray(path, reduce)


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
_CPU = object()          # marks a photo the decoder left to the CPU fallback


def _clahe():
    """The worker's CLAHE object (clip 2.0, 8 x 8 tiles, features.py:70), created on first use."""
    if not _CLAHE:
        from .clahe import createCLAHE
        _CLAHE.append(createCLAHE(2.0, (8, 8)))
    return _CLAHE[0]


def gpu_worker_main(task_q, res_q, n_threads, wid, batch=16, decode=read_gray, clahe=False, decode_task=None,
                    cpu_fallback=None):
    """features.py:270-381 with the one-photo GPU loop replaced by batches.  ``n_threads``
    decoder threads take tasks (idx, path, out_npz, thumb_path, reduce, nfeat) from task_q
    (None ends a thread) and decode them; the GPU loop gathers what is ready (at least one,
    at most ``batch`` photos), runs them through one FeatureBatch per nfeat (the reference's
    sift_gpu settings: max_kp = 2.2 nfeat, contrast 0.015, edge 12), writes each feature file
    with _save and posts ("ok", (idx, n, err, timings)).  Also posted: ("gpu", wid, info) when
    ready, ("gpu_off", wid, msg) if the GPU cannot be used (without cpu_fallback every photo is
    then reported failed), and ("exit", "gpu%d" % wid) at the end.  The default decoder,
    read_gray, reads .npy gray images and JPEG photos reduced by the task's reduce; another
    ``decode`` is called with the path alone.  With ``clahe=True`` the decoder returns the plain
    gray image and the contrast equalisation of features.py:70-71 (clip 2.0, 8 x 8 tiles) runs
    on the GPU on each photo of the batch.  decode_task (optional): called with the whole task
    tuple instead, for a decoder that needs reduce / thumb_path: prepare_task is the
    reference's prepare (features.py:60-71) with its CLAHE left to ``clahe=True``.
    cpu_fallback (optional): the CPU path of features.py:355-364 - called with a task tuple, it
    returns (idx, n, err, timings) (the reference's features.extract); when given, the photos go
    to it while the GPU is unavailable ("gpu_off") or after a GPU batch fails (then "gpu_off" is
    posted once and the rest of the photos stay on the CPU).  Both must be picklable when the
    worker is a spawned process (module-level functions)."""
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
    gpu_state = {"off": gpu_err is not None}

    def decoder():
        while True:
            task = task_q.get()
            if task is None:
                ready.put(None)
                break
            tt = {"read": 0.0, "decode": 0.0, "bytes": 0, "photos_gpu": 1}
            if gpu_state["off"] and cpu_fallback is not None:
                ready.put((task, None, tt, _CPU))       # decoded by the CPU path itself
                continue
            try:
                t1 = time.perf_counter()
                if decode_task is not None:
                    g2 = decode_task(task)
                elif decode is read_gray:
                    g2 = read_gray(task[1], int(task[4]))      # the task's reduce
                else:
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
            size, angle = fb.size_angle(k)
            n = _save(task[2], xy, None, resp, g2.shape[1], g2.shape[0], desc_u8=u8, size=size, angle=angle)
            tt["gpu_detect"] = (time.perf_counter() - t2) / len(group)
            tt["gpu_describe"] = 0.0
            res_q.put(("ok", (task[0], n, None, tt)))

    def on_cpu(task):
        res_q.put(("ok", cpu_fallback(task)))

    def process(items):
        good = []
        for task, g2, tt, err in items:
            if err is _CPU or (gpu_state["off"] and cpu_fallback is not None and not err):
                on_cpu(task)
            elif err or gpu_state["off"]:
                res_q.put(("ok", (task[0], 0, err or ("GPU unavailable: " + str(gpu_err)), tt)))
            else:
                good.append((task, g2, tt, None))
        by_nfeat = {}
        for it in good:
            by_nfeat.setdefault(int(it[0][5]), []).append(it)
        for group in by_nfeat.values():
            if gpu_state["off"] and cpu_fallback is not None:
                for task, _, _, _ in group:
                    on_cpu(task)
                continue
            try:
                run_group(group)
            except Exception as e:  # noqa: BLE001 - the GPU failed on this batch
                msg = "%s: %s" % (type(e).__name__, e)
                if cpu_fallback is None:
                    for task, _, tt, _ in group:
                        res_q.put(("ok", (task[0], 0, msg, tt)))
                else:
                    gpu_state["off"] = True
                    res_q.put(("gpu_off", wid, "GPU failed during processing, continuing on the CPU: %s" % msg[:300]))
                    for task, _, _, _ in group:
                        on_cpu(task)

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
