"""Learned match filter (no counterpart in the reference): a 264-256-256-1 MLP with ReLU and a
sigmoid output that scores a putative match (a query and its best candidate in the prior
window) as a likely geometric inlier.  Weights and activations are BF16, products accumulate
in FP32 on MFMA; training is binary cross-entropy with Adam on FP32 master weights.  Every
step of the forward and backward pass runs in libfsx's own kernels (mlp.hip); torch only
holds the buffers.

Inputs are padded from 264 to 288 columns (a multiple of the GEMM's K step); the pad columns
and the matching rows of W1 stay zero.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

N_IN, N_INP, N_HID = 264, 288, 256
WGRAD_SPLITS = 16          # K chunks of the weight-gradient products (K = batch rows)


class MatchFilterMLP:
    def __init__(self, seed=0, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, device="cuda"):
        import torch
        self.torch = torch
        self.dev = device
        self.lr, self.b1_, self.b2_, self.eps = float(lr), float(betas[0]), float(betas[1]), float(eps)
        rng = np.random.default_rng(seed)
        W1 = np.zeros((N_INP, N_HID), np.float32)
        W1[:N_IN] = rng.normal(0.0, np.sqrt(2.0 / N_IN), (N_IN, N_HID))
        W2 = rng.normal(0.0, np.sqrt(2.0 / N_HID), (N_HID, N_HID)).astype(np.float32)
        w3 = rng.normal(0.0, np.sqrt(1.0 / N_HID), N_HID).astype(np.float32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device)   # noqa: E731
        self.p = {"W1": t(W1), "b1": t(np.zeros(N_HID)), "W2": t(W2), "b2": t(np.zeros(N_HID)),
                  "w3": t(w3), "b3": t(np.zeros(1))}
        self.m = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.v = {k: torch.zeros_like(v) for k, v in self.p.items()}
        self.W1_16 = self.p["W1"].to(torch.bfloat16)
        self.W2_16 = self.p["W2"].to(torch.bfloat16)
        self.step_count = 0
        self._bufs = {}

    # -------------------------------------------------------------------------------------
    def _buf(self, nb):
        torch = self.torch
        b = self._bufs.get(nb)
        if b is None:
            bf = torch.bfloat16
            sc = C.c_size_t()
            _lib.call("fsx_mlp_colsum_scratch", nb, N_HID + 2, C.byref(sc))
            sk = C.c_size_t()
            _lib.call("fsx_mlp_splitk_scratch", N_INP, N_HID, WGRAD_SPLITS, C.byref(sk))
            b = {"H1": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "H2": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "dZ2": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "dZ1": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "p": torch.empty(nb, dtype=torch.float32, device=self.dev),
                 "part": torch.empty(((nb + 3) // 4, N_HID + 2), dtype=torch.float32, device=self.dev),
                 "hp": torch.empty(N_HID + 2, dtype=torch.float32, device=self.dev),
                 "scratch": torch.empty(max(sc.value // 4, 1), dtype=torch.float32, device=self.dev),
                 "splitk": torch.empty(max(sk.value // 4, 1), dtype=torch.float32, device=self.dev),
                 "g": {"W1": torch.empty((N_INP, N_HID), dtype=torch.float32, device=self.dev),
                       "W2": torch.empty((N_HID, N_HID), dtype=torch.float32, device=self.dev),
                       "b1": torch.empty(N_HID, dtype=torch.float32, device=self.dev),
                       "b2": torch.empty(N_HID, dtype=torch.float32, device=self.dev)}}
            self._bufs = {nb: b}
        return b

    @staticmethod
    def _P(t):
        return C.c_void_p(t.data_ptr())

    def _st(self):
        return C.c_void_p(_lib.stream_handle())

    def _forward(self, X16, b):
        nb = X16.shape[0]
        P = self._P
        _lib.call("fsx_mlp_gemm", 0, 0, 0, P(X16), N_INP, P(self.W1_16), N_HID, nb, N_HID, N_INP, None,
                  P(b["H1"]), N_HID, P(self.p["b1"]), None, self._st())
        _lib.call("fsx_mlp_gemm", 0, 0, 0, P(b["H1"]), N_HID, P(self.W2_16), N_HID, nb, N_HID, N_HID, None,
                  P(b["H2"]), N_HID, P(self.p["b2"]), None, self._st())

    def forward(self, X16):
        """Probabilities for the rows of X16 (bf16 [n, 288] device tensor)."""
        self._check(X16)
        b = self._buf(X16.shape[0])
        self._forward(X16, b)
        P = self._P
        _lib.call("fsx_mlp_head", P(b["H2"]), P(self.p["w3"]), P(self.p["b3"]), None, X16.shape[0], P(b["p"]),
                  None, None, self._st())
        return b["p"]

    def gradients(self, X16, y):
        """Forward + backward on one batch; returns (loss, dict of float32 gradients)."""
        self._check(X16)
        nb = X16.shape[0]
        if nb % 32:
            raise ValueError("batch size must be a multiple of 32")
        b = self._buf(nb)
        P = self._P
        self._forward(X16, b)
        _lib.call("fsx_mlp_head", P(b["H2"]), P(self.p["w3"]), P(self.p["b3"]), P(y), nb, P(b["p"]), P(b["dZ2"]),
                  P(b["part"]), self._st())
        _lib.call("fsx_mlp_colsum", P(b["part"]), 0, b["part"].shape[0], N_HID + 2, N_HID + 2, P(b["hp"]),
                  P(b["scratch"]), self._st())
        g = b["g"]
        # dW2 = H1^T dZ2 ; db2 = colsum(dZ2) ; dZ1 = (dZ2 W2^T) masked by H1 > 0
        _lib.call("fsx_mlp_gemm_splitk", 1, 0, P(b["H1"]), N_HID, P(b["dZ2"]), N_HID, N_HID, N_HID, nb, P(g["W2"]),
                  N_HID, P(b["splitk"]), WGRAD_SPLITS, self._st())
        _lib.call("fsx_mlp_colsum", P(b["dZ2"]), 1, nb, N_HID, N_HID, P(g["b2"]), P(b["scratch"]), self._st())
        _lib.call("fsx_mlp_gemm", 0, 1, 2, P(b["dZ2"]), N_HID, P(self.W2_16), N_HID, nb, N_HID, N_HID, None,
                  P(b["dZ1"]), N_HID, None, P(b["H1"]), self._st())
        # dW1 = X^T dZ1 ; db1 = colsum(dZ1)
        _lib.call("fsx_mlp_gemm_splitk", 1, 0, P(X16), N_INP, P(b["dZ1"]), N_HID, N_INP, N_HID, nb, P(g["W1"]),
                  N_HID, P(b["splitk"]), WGRAD_SPLITS, self._st())
        _lib.call("fsx_mlp_colsum", P(b["dZ1"]), 1, nb, N_HID, N_HID, P(g["b1"]), P(b["scratch"]), self._st())
        hp = b["hp"]
        grads = {"W1": g["W1"], "b1": g["b1"], "W2": g["W2"], "b2": g["b2"], "w3": hp[:N_HID], "b3": hp[N_HID:N_HID + 1]}
        return hp[N_HID + 1] / nb, grads

    def step(self, X16, y):
        """One Adam step on a batch; returns the batch loss (device scalar)."""
        loss, grads = self.gradients(X16, y)
        self.step_count += 1
        P = self._P
        for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
            w16 = self.W1_16 if k == "W1" else (self.W2_16 if k == "W2" else None)
            _lib.call("fsx_mlp_adam", P(self.p[k]), P(self.m[k]), P(self.v[k]), P(grads[k].contiguous()),
                      P(w16) if w16 is not None else None, self.p[k].numel(), self.lr, self.b1_, self.b2_, self.eps,
                      self.step_count, self._st())
        return loss

    def predict(self, X16, chunk=None):
        """Probabilities for any number of rows: one fused launch (fsx_mlp_infer) over the
        weights packed for it, packed again whenever an Adam step has changed them.  chunk is
        accepted for compatibility and ignored."""
        torch = self.torch
        self._check(X16)
        out = torch.empty(X16.shape[0], dtype=torch.float32, device=self.dev)
        if X16.shape[0] == 0:
            return out
        if getattr(self, "_packed_at", None) != self.step_count:
            nb = C.c_size_t()
            _lib.call("fsx_mlp_infer_pack_bytes", C.byref(nb))
            self._packed = torch.empty(nb.value // 2, dtype=torch.int16, device=self.dev)
            _lib.call("fsx_mlp_infer_pack", self._P(self.W1_16), self._P(self.W2_16), self._P(self._packed), self._st())
            self._packed_at = self.step_count
        P = self._P
        _lib.call("fsx_mlp_infer", P(X16), X16.shape[0], P(self._packed), P(self.p["b1"]), P(self.p["b2"]),
                  P(self.p["w3"]), P(self.p["b3"]), P(out), self._st())
        return out

    def _check(self, X16):
        torch = self.torch
        if X16.dtype != torch.bfloat16 or X16.dim() != 2 or X16.shape[1] != N_INP or not X16.is_contiguous():
            raise ValueError("X16 must be a contiguous bfloat16 [n, %d] tensor" % N_INP)


# ------------------------------------------------------------------------------------------
# Training data from the plain pipeline and the filter's evaluation

def flight_candidates(seed, workers=8, radius=60.0, ratio=0.8, n_hyp=512, thr=3.0, nfeat=8000, chunk=32,
                      keep_matcher=False, **flight_kw):
    """Candidates and labels of one synthetic flight (synth.flight with flight_kw: a 2 x 4 grid
    by default, 20 x 20 for the 400-photo flights): SIFT on the GPU in chunks of `chunk` photos,
    guided matching of every overlapping pair (the plain method), homography RANSAC on each
    pair's plain matches.  Returns a dict with X (bf16 [n, 288] device rows, one per query with
    a candidate), y (float32 device labels: 1 = the plain pipeline's match and a RANSAC
    inlier), plain (bool numpy: accepted by the ratio test and the one-query-per-target rule)
    and counts.  keep_matcher=True adds the flight's PairMatcher ("matcher", after its plain run),
    the per-photo keypoint sizes / angles and positions, the labelling RANSAC's settings and
    per-query arrays for checking the pruning stage."""
    import torch
    from . import features, synth
    from .matching import PairMatcher
    from .verify import ransac_homography
    fl = synth.flight(seed=seed, workers=workers, **flight_kw)
    shapes = [im.shape for im in fl.images]
    n_img = len(shapes)
    images, sizes, angles, xys = [None] * n_img, [None] * n_img, [None] * n_img, [None] * n_img
    for c0 in range(0, n_img, chunk):
        idx = list(range(c0, min(n_img, c0 + chunk)))
        fb = features.FeatureBatch([shapes[i] for i in idx], nfeat=nfeat, contrast=0.04, edge=10.0)
        fb.run([torch.from_numpy(fl.images[i]).cuda() for i in idx])
        fb.check_capacity()
        for k, i in enumerate(idx):
            H, W = shapes[i]
            xy, u8, _, _ = fb.results(k)
            sel = fb._sel(k, len(xy))
            kp = fb.kp[k].cpu().numpy()
            xy64 = xy.astype(np.float64)
            images[i] = (fb.u8[k, :len(xy)].clone(), torch.from_numpy(xy64).cuda(), (W, H))
            sizes[i] = kp[2][sel]
            angles[i] = kp[3][sel]
            xys[i] = xy64
        del fb
    pairs = list(fl.pairs)
    pm = PairMatcher(images, pairs, [fl.prior[p] for p in pairs], radius=radius, kc=0, ratio=ratio)
    pm.run("mfma")
    res = pm.results()
    Xall = pm.filter_features(sizes, angles)
    cbest, _, _ = pm.candidates()
    cb = cbest.cpu().numpy()
    qoff = pm.query_offsets()
    # RANSAC labels on each pair's plain matches
    pa = [xys[i][q] for (i, j), (q, t) in zip(pairs, res)]
    pb = [xys[j][t] for (i, j), (q, t) in zip(pairs, res)]
    flags, _, _ = ransac_homography(pa, pb, n_hyp=n_hyp, thr=thr, seed=0x5EED + seed)
    nq = len(cb)
    label = np.zeros(nq, np.float32)
    plain = np.zeros(nq, bool)
    for p, ((q, t), f) in enumerate(zip(res, flags)):
        plain[qoff[p] + q] = True
        label[qoff[p] + q[f]] = 1.0
    keep = cb >= 0
    idx = torch.from_numpy(np.nonzero(keep)[0]).cuda()
    out = {"X": Xall.index_select(0, idx).contiguous(), "y": torch.from_numpy(label[keep]).cuda(),
           "plain": plain[keep], "label": label[keep] > 0.5, "n_query": nq, "n_pairs": len(pairs),
           "n_plain": int(plain.sum()), "n_inliers": int(label.sum())}
    if keep_matcher:
        out.update({"matcher": pm, "sizes": sizes, "angles": angles, "query_offsets": qoff, "plain_results": res,
                    "query_label": label > 0.5, "xy": xys, "ransac": (n_hyp, thr, 0x5EED + seed)})
    return out


def train(net, X, y, epochs=4, batch=8192, seed=0):
    """Adam epochs over (X, y) in seeded shuffled batches (the tail batch is dropped);
    returns the mean loss of each epoch."""
    import torch
    n = X.shape[0]
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(epochs):
        perm = torch.from_numpy(rng.permutation(n)).cuda()
        losses = []
        for s in range(0, n - batch + 1, batch):
            ib = perm[s:s + batch]
            losses.append(net.step(X.index_select(0, ib).contiguous(), y.index_select(0, ib).contiguous()))
        out.append(float(torch.stack(losses).mean()))
    return out


def threshold_for_recall(p, positive, recall=0.995):
    """Largest score threshold keeping at least `recall` of the positives."""
    s = np.sort(np.asarray(p)[np.asarray(positive)])
    if len(s) == 0:
        return 0.5
    k = int(np.floor((1.0 - recall) * len(s)))
    return float(s[k])


def evaluate(p, tau, data):
    """Recall / precision of the pruned candidate set against the plain pipeline's verified
    inliers, and what survives into verification with and without the network."""
    p = np.asarray(p)
    kept = p >= tau
    pos = data["label"]
    plain = data["plain"]
    n = len(p)
    both = kept & plain
    return {"candidates": n, "kept": int(kept.sum()), "pruned_fraction": float(1.0 - kept.mean()),
            "inlier_recall": float((kept & pos).sum() / max(pos.sum(), 1)),
            "precision_kept": float((kept & pos).sum() / max(kept.sum(), 1)),
            "plain_ratio_test": {"accepted": int(plain.sum()), "inlier_fraction": float(pos.sum() / max(plain.sum(), 1))},
            "filter_then_ratio_test": {"accepted": int(both.sum()),
                                       "inlier_fraction": float((both & pos).sum() / max(both.sum(), 1)),
                                       "inlier_recall": float((both & pos).sum() / max(pos.sum(), 1))}}
