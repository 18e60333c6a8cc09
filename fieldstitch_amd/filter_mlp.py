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
            b = {"H1": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "H2": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "dZ2": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "dZ1": torch.empty((nb, N_HID), dtype=bf, device=self.dev),
                 "p": torch.empty(nb, dtype=torch.float32, device=self.dev),
                 "part": torch.empty(((nb + 3) // 4, N_HID + 2), dtype=torch.float32, device=self.dev),
                 "hp": torch.empty(N_HID + 2, dtype=torch.float32, device=self.dev),
                 "scratch": torch.empty(max(sc.value // 4, 1), dtype=torch.float32, device=self.dev),
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
        _lib.call("fsx_mlp_gemm", 1, 0, 1, P(b["H1"]), N_HID, P(b["dZ2"]), N_HID, N_HID, N_HID, nb, P(g["W2"]),
                  None, N_HID, None, None, self._st())
        _lib.call("fsx_mlp_colsum", P(b["dZ2"]), 1, nb, N_HID, N_HID, P(g["b2"]), P(b["scratch"]), self._st())
        _lib.call("fsx_mlp_gemm", 0, 1, 2, P(b["dZ2"]), N_HID, P(self.W2_16), N_HID, nb, N_HID, N_HID, None,
                  P(b["dZ1"]), N_HID, None, P(b["H1"]), self._st())
        # dW1 = X^T dZ1 ; db1 = colsum(dZ1)
        _lib.call("fsx_mlp_gemm", 1, 0, 1, P(X16), N_INP, P(b["dZ1"]), N_HID, N_INP, N_HID, nb, P(g["W1"]),
                  None, N_HID, None, None, self._st())
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

    def predict(self, X16, chunk=65536):
        """Probabilities for any number of rows (chunked)."""
        torch = self.torch
        out = torch.empty(X16.shape[0], dtype=torch.float32, device=self.dev)
        for s in range(0, X16.shape[0], chunk):
            out[s:s + chunk] = self.forward(X16[s:s + chunk])
        return out

    def _check(self, X16):
        torch = self.torch
        if X16.dtype != torch.bfloat16 or X16.dim() != 2 or X16.shape[1] != N_INP or not X16.is_contiguous():
            raise ValueError("X16 must be a contiguous bfloat16 [n, %d] tensor" % N_INP)
