"""CPU restatement of the deterministic homography RANSAC (test infrastructure only): the same
counter-hash sampling, 4-point DLT solve (Gaussian elimination with partial pivoting, float64,
no contraction) and inlier test as fieldstitch_amd/csrc/verify.hip, vectorised over the
hypotheses, so every count and label can be compared exactly."""
from __future__ import annotations

import numpy as np

M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x):
    with np.errstate(over="ignore"):
        x = np.asarray(x, np.uint64) + np.uint64(0x9E3779B97F4A7C15)
        x = (x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        x = (x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return x ^ (x >> np.uint64(31))


def draw(seed, p, h, j):
    inner = (np.uint64(p) << np.uint64(40)) ^ (np.asarray(h, np.uint64) << np.uint64(8)) ^ np.uint64(j)
    return splitmix64(np.uint64(seed) ^ splitmix64(inner))


def samples(seed, p, n, hyps):
    """(n_hyp, 4) indices and a validity mask."""
    hyps = np.asarray(hyps, np.int64)
    idx = np.full((len(hyps), 4), -1, np.int64)
    ok = np.ones(len(hyps), bool)
    for j in range(4):
        cand = np.full(len(hyps), -1, np.int64)
        for attempt in range(16):
            c = (draw(seed, p, hyps, j * 16 + attempt) % np.uint64(n)).astype(np.int64)
            dup = np.zeros(len(hyps), bool)
            for u in range(j):
                dup |= idx[:, u] == c
            take = (cand < 0) & ~dup
            cand = np.where(take, c, cand)
        ok &= cand >= 0
        idx[:, j] = cand
    return idx, ok


def solve_h4(xa, ya, xb, yb):
    """Batched 4-point DLT: arrays (m, 4) -> (h (m, 9), valid (m,))."""
    m = xa.shape[0]
    A = np.zeros((m, 8, 9))
    for k in range(4):
        x, y, u, v = xa[:, k], ya[:, k], xb[:, k], yb[:, k]
        A[:, 2 * k, 0], A[:, 2 * k, 1], A[:, 2 * k, 2] = x, y, 1.0
        A[:, 2 * k, 6], A[:, 2 * k, 7], A[:, 2 * k, 8] = -(u * x), -(u * y), u
        A[:, 2 * k + 1, 3], A[:, 2 * k + 1, 4], A[:, 2 * k + 1, 5] = x, y, 1.0
        A[:, 2 * k + 1, 6], A[:, 2 * k + 1, 7], A[:, 2 * k + 1, 8] = -(v * x), -(v * y), v
    valid = np.ones(m, bool)
    r_ = np.arange(m)
    for c in range(8):
        col = np.abs(A[:, c:, c])
        piv = c + np.argmax(col, axis=1)
        best = col[r_, piv - c]
        valid &= best >= 1e-12
        rowc = A[r_, c].copy()
        A[r_, c] = A[r_, piv]
        A[r_, piv] = rowc
        for r in range(c + 1, 8):
            with np.errstate(divide="ignore", invalid="ignore"):
                f = A[:, r, c] / A[:, c, c]
            for k in range(c, 9):
                A[:, r, k] = A[:, r, k] - f * A[:, c, k]
    h = np.zeros((m, 9))
    with np.errstate(divide="ignore", invalid="ignore"):
        for c in range(7, -1, -1):
            s = A[:, c, 8].copy()
            for k in range(c + 1, 8):
                s = s - A[:, c, k] * h[:, k]
            h[:, c] = s / A[:, c, c]
    h[:, 8] = 1.0
    return h, valid


def inliers(h, a, b, thr2):
    """h (m, 9); a, b (n, 2) -> (m, n) bool."""
    x, y, u, v = a[None, :, 0], a[None, :, 1], b[None, :, 0], b[None, :, 1]
    H = [h[:, k:k + 1] for k in range(9)]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        w = (H[6] * x + H[7] * y) + H[8]
        good = np.abs(w) > 1e-12
        px = ((H[0] * x + H[1] * y) + H[2]) / w
        py = ((H[3] * x + H[4] * y) + H[5]) / w
    dx, dy = px - u, py - v
    return good & (dx * dx + dy * dy < thr2)


def ransac_pair(p, a, b, n_hyp, thr, seed):
    """Returns (inlier flags (n,), H (9,), count) for pair index p."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = len(a)
    if n < 4:
        return np.zeros(n, bool), np.zeros(9), 0
    hyps = np.arange(n_hyp)
    idx, ok = samples(seed, p, n, hyps)
    safe = np.where(ok[:, None], idx, 0)
    h, valid = solve_h4(a[safe, 0], a[safe, 1], b[safe, 0], b[safe, 1])
    valid &= ok
    cnt = np.where(valid, inliers(h, a, b, thr * thr).sum(1), -1)
    if cnt.max() < 0:
        return np.zeros(n, bool), np.zeros(9), 0
    best = int(np.argmax(cnt))          # first maximum = lowest hypothesis index
    return inliers(h[best:best + 1], a, b, thr * thr)[0], h[best], int(cnt[best])
