"""Feature tracks from verified match lists on the GPU (replaces recon.py:14-56, build_tracks).

Nodes are the features of all photos numbered photo by photo (offsets = exclusive cumulative
feature counts); each verified match (i, j, q, t) joins offsets[i] + q and offsets[j] + t.  A
track is a connected component of the matched features with at least two members and at most
one feature per photo (components holding two features of one photo are dropped, as in
recon.py:37-48).  Tracks are numbered in increasing order of their smallest node and the
observations come track by track, nodes increasing (recon.py:49-56).  The components come from a
lock-free union-find in libfsx (`fsx_build_tracks`); the outputs equal the reference's
int64 arrays (tid, img, feat) and its track count.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class TrackBuilder:
    """Reusable device workspace for build_tracks on device-resident match lists."""

    def __init__(self, device="cuda"):
        self.device = device
        self._ws = None
        self._ws_bytes = 0
        self._out = None

    def _workspace(self, n_nodes, n_edges):
        import torch
        nb = C.c_size_t()
        cap = C.c_longlong()
        _lib.call("fsx_tracks_workspace_bytes", int(n_nodes), int(n_edges), C.byref(nb), C.byref(cap))
        if self._ws is None or self._ws_bytes < nb.value:
            self._ws = torch.empty(max(1, nb.value), dtype=torch.uint8, device=self.device)
            self._ws_bytes = nb.value
        c = max(1, cap.value)
        if self._out is None or self._out.shape[1] < c:
            self._out = torch.empty((3, c), dtype=torch.int64, device=self.device)
        return cap.value

    def run(self, offsets_d, pair_ij_d, pair_start_d, q_d, t_d, stream=None):
        """Device form.  offsets_d: (n_img + 1,) int64; pair_ij_d: (n_pair, 2) int32;
        pair_start_d: (n_pair + 1,) int64; q_d, t_d: (n_edges,) int32, all on the GPU.
        Returns (tid, img, feat) int64 device views of length n_observations, and n_tracks;
        synchronises the stream once to read the two counts.  With `stream`, every buffer,
        launch and copy of the call is ordered on it."""
        import torch
        from contextlib import nullcontext
        with torch.cuda.stream(stream) if stream is not None else nullcontext():
            return self._run(offsets_d, pair_ij_d, pair_start_d, q_d, t_d, stream)

    def _run(self, offsets_d, pair_ij_d, pair_start_d, q_d, t_d, stream):
        import torch
        n_img = int(offsets_d.shape[0]) - 1
        n_pair = int(pair_ij_d.shape[0])
        n_edges = int(q_d.shape[0])
        if n_img < 1:
            raise ValueError("build_tracks: no photos")
        if t_d.shape[0] != n_edges or pair_start_d.shape[0] != n_pair + 1:
            raise ValueError("build_tracks: inconsistent match list shapes")
        for x, dt in ((offsets_d, torch.int64), (pair_ij_d, torch.int32), (pair_start_d, torch.int64),
                      (q_d, torch.int32), (t_d, torch.int32)):
            if x.dtype != dt or not x.is_cuda or not x.is_contiguous():
                raise TypeError("build_tracks: expected contiguous %s device tensors" % dt)
        n_nodes = int(offsets_d[-1].item()) if n_img > 0 else 0
        cap = self._workspace(n_nodes, n_edges)
        res = torch.empty(4, dtype=torch.int64, device=self.device)
        o = self._out
        vp = lambda x: C.c_void_p(x.data_ptr())
        _lib.call("fsx_build_tracks", n_img, vp(offsets_d), n_nodes, n_pair, vp(pair_ij_d), vp(pair_start_d),
                  vp(q_d), vp(t_d), n_edges, vp(self._ws), self._ws_bytes, o.shape[1], vp(o[0]), vp(o[1]),
                  vp(o[2]), vp(res), C.c_void_p(_lib.stream_handle(stream)))
        r = res.cpu().tolist()
        if r[2]:
            raise ValueError("build_tracks: an image or feature index is out of range")
        if r[3] or r[0] > cap:
            raise RuntimeError("build_tracks: observation capacity exceeded")
        n_obs = int(r[0])
        return o[0, :n_obs], o[1, :n_obs], o[2, :n_obs], int(r[1])


def pack_pairs(pairs, nfeat):
    """Host packing of the reference's match list [(i, j, q, t), ...] (recon.py:20-23): int64
    offsets, (n_pair, 2) int32 images, int64 pair starts, int32 q / t."""
    nfeat = np.asarray(nfeat, np.int64)
    offsets = np.concatenate([[0], np.cumsum(nfeat)]).astype(np.int64)
    if offsets[-1] >= 2 ** 31:
        raise ValueError("build_tracks: %d features in all; at most 2^31 - 1" % offsets[-1])
    ij = np.array([(int(i), int(j)) for i, j, _, _ in pairs], np.int32).reshape(-1, 2)
    lens = np.array([len(q) for _, _, q, _ in pairs], np.int64)
    for _, _, q, t in pairs:
        if len(q) != len(t):
            raise ValueError("build_tracks: q and t differ in length")
    start = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    q = np.concatenate([np.asarray(q, np.int64) for _, _, q, _ in pairs]) if pairs else np.zeros(0, np.int64)
    t = np.concatenate([np.asarray(t, np.int64) for _, _, _, t in pairs]) if pairs else np.zeros(0, np.int64)
    if len(q) and (q.min() < 0 or t.min() < 0 or q.max() >= 2 ** 31 or t.max() >= 2 ** 31):
        raise ValueError("build_tracks: a feature index is out of range")
    return offsets, ij, start, q.astype(np.int32), t.astype(np.int32)


_BUILDER = {}


def build_tracks(pairs, nfeat, n_img, stream=None):
    """recon.py:14-56.  pairs: list of (i, j, qi, tj); nfeat: features per photo; n_img: photos
    (>= len(nfeat)).  Returns (tid, img, feat) int64 numpy arrays (observations sorted by track,
    then photo and feature) and the number of tracks."""
    import torch
    if not pairs:
        z = np.zeros(0, np.int64)
        return z, z.copy(), z.copy(), 0
    nfeat = np.asarray(nfeat, np.int64)
    if int(n_img) < len(nfeat):
        raise ValueError("build_tracks: n_img=%d < %d photos with features" % (n_img, len(nfeat)))
    offsets, ij, start, q, t = pack_pairs(pairs, nfeat)
    dev = torch.device("cuda", torch.cuda.current_device())
    b = _BUILDER.get(dev.index)
    if b is None:
        b = _BUILDER[dev.index] = TrackBuilder(dev)
    d = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (offsets, ij, start, q, t)]
    tid, img, feat, n_tracks = b.run(*d, stream=stream)
    return tid.cpu().numpy(), img.cpu().numpy(), feat.cpu().numpy(), n_tracks
