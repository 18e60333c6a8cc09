"""CPU restatement of the deterministic RANSAC (test infrastructure only): the same
counter-hash sampling, model solves (4-point homography DLT with partial pivoting; normalized
8-point fundamental matrix with full pivoting and a Jacobi rank-2 projection; 2-point
similarity; float64, no contraction, sums in the device's order), inlier tests (fused
multiply-adds where the device fuses them, through the C fma of sift_oracle.c) and
least-squares refit of the winning homography as fieldstitch_amd/csrc/verify.hip, vectorised
over the hypotheses, so every count, label and model can be compared exactly; and the adaptive
stopping of fsx_ransac_batch_adaptive (rounds of 256 hypotheses, the stopping rule in the same
float64 multiplications), pinned against the closed form log(1 - conf) / log(1 - w^K)."""
from __future__ import annotations

import numpy as np

from oracle.sift_oracle import fma

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


def samples(seed, p, n, hyps, K=4):
    """(n_hyp, K) indices and a validity mask."""
    hyps = np.asarray(hyps, np.int64)
    idx = np.full((len(hyps), K), -1, np.int64)
    ok = np.ones(len(hyps), bool)
    for j in range(K):
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
    """h (m, 9); a, b (n, 2) -> (m, n) bool: reprojection error below sqrt(thr2), tested as
    |H p - w q|^2 < thr2 w^2 (no division), with verify.hip's fused multiply-adds:
    w = fma(h6, x, fma(h7, y, h8)), ex = fma(-u, w, fma(h0, x, fma(h1, y, h2))), ey likewise,
    |e|^2 = fma(ex, ex, ey * ey)."""
    x, y, u, v = a[None, :, 0], a[None, :, 1], b[None, :, 0], b[None, :, 1]
    H = [h[:, k:k + 1] for k in range(9)]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        w = fma(H[6], x, fma(H[7], y, H[8]))
        good = np.abs(w) > 1e-12
        ex = fma(-u, w, fma(H[0], x, fma(H[1], y, H[2])))
        ey = fma(-v, w, fma(H[3], x, fma(H[4], y, H[5])))
        return good & (fma(ex, ex, ey * ey) < thr2 * (w * w))


def rank2_project(f):
    """Batched closest rank-2 matrix (verify.hip rank2_project): f (m, 9) row-major, scaled by
    its largest |entry|, then f (I - v v^T) with v the smallest-eigenvalue eigenvector of f^T f
    from six cyclic Jacobi sweeps (rotations (0,1), (0,2), (1,2)); rows with f = 0 unchanged."""
    f = np.array(f, np.float64)
    m = f.shape[0]
    mx = np.abs(f).max(axis=1)
    live = mx > 0.0
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        f = np.where(live[:, None], f / np.where(live, mx, 1.0)[:, None], f)
        S = np.zeros((m, 3, 3))
        V = np.zeros((m, 3, 3))
        for i in range(3):
            for j in range(3):
                S[:, i, j] = (f[:, i] * f[:, j] + f[:, 3 + i] * f[:, 3 + j]) + f[:, 6 + i] * f[:, 6 + j]
            V[:, i, i] = 1.0
        for _ in range(6):
            for p, q in ((0, 1), (0, 2), (1, 2)):
                apq = S[:, p, q].copy()
                rot = apq != 0.0
                theta = (S[:, q, q] - S[:, p, p]) / (2.0 * np.where(rot, apq, 1.0))
                t = np.where(theta >= 0.0, 1.0, -1.0) / (np.abs(theta) + np.sqrt(theta * theta + 1.0))
                c = 1.0 / np.sqrt(t * t + 1.0)
                sn = t * c
                c = np.where(rot, c, 1.0)
                sn = np.where(rot, sn, 0.0)
                S0, V0 = S.copy(), V.copy()
                for k in range(3):
                    S[:, k, p] = np.where(rot, c * S0[:, k, p] - sn * S0[:, k, q], S0[:, k, p])
                    S[:, k, q] = np.where(rot, sn * S0[:, k, p] + c * S0[:, k, q], S0[:, k, q])
                    V[:, k, p] = np.where(rot, c * V0[:, k, p] - sn * V0[:, k, q], V0[:, k, p])
                    V[:, k, q] = np.where(rot, sn * V0[:, k, p] + c * V0[:, k, q], V0[:, k, q])
                S1 = S.copy()
                for k in range(3):
                    S[:, p, k] = np.where(rot, c * S1[:, p, k] - sn * S1[:, q, k], S1[:, p, k])
                    S[:, q, k] = np.where(rot, sn * S1[:, p, k] + c * S1[:, q, k], S1[:, q, k])
        d = np.stack([S[:, 0, 0], S[:, 1, 1], S[:, 2, 2]], axis=1)
        kmin = np.zeros(m, np.int64)
        for k in (1, 2):
            kmin = np.where(d[:, k] < d[np.arange(m), kmin], k, kmin)
        v = V[np.arange(m), :, kmin]
        out = f.copy()
        for i in range(3):
            w = (f[:, 3 * i] * v[:, 0] + f[:, 3 * i + 1] * v[:, 1]) + f[:, 3 * i + 2] * v[:, 2]
            for j in range(3):
                out[:, 3 * i + j] = f[:, 3 * i + j] - w * v[:, j]
    return np.where(live[:, None], out, f)


def solve_f8(xa, ya, xb, yb):
    """Batched normalized 8-point fundamental matrix: arrays (m, 8) -> (F (m, 9), valid (m,)).
    Centroid / mean-distance normalization per image, full-pivot elimination of the 8 x 9
    system, free unknown = 1, projected to rank 2, F = Tb^T F^ Ta (verify.hip ModelF)."""
    m = xa.shape[0]
    r_ = np.arange(m)
    ca = [np.zeros(m), np.zeros(m)]
    cb = [np.zeros(m), np.zeros(m)]
    for k in range(8):
        ca[0] = ca[0] + xa[:, k]
        ca[1] = ca[1] + ya[:, k]
        cb[0] = cb[0] + xb[:, k]
        cb[1] = cb[1] + yb[:, k]
    ca = [c / 8.0 for c in ca]
    cb = [c / 8.0 for c in cb]
    ra = np.zeros(m)
    rb = np.zeros(m)
    for k in range(8):
        ra = ra + np.sqrt((xa[:, k] - ca[0]) * (xa[:, k] - ca[0]) + (ya[:, k] - ca[1]) * (ya[:, k] - ca[1]))
        rb = rb + np.sqrt((xb[:, k] - cb[0]) * (xb[:, k] - cb[0]) + (yb[:, k] - cb[1]) * (yb[:, k] - cb[1]))
    valid = (ra > 1e-12) & (rb > 1e-12)
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        sa = 1.4142135623730951 * 8.0 / ra
        sb = 1.4142135623730951 * 8.0 / rb
        A = np.zeros((m, 8, 9))
        for k in range(8):
            x = (xa[:, k] - ca[0]) * sa
            y = (ya[:, k] - ca[1]) * sa
            u = (xb[:, k] - cb[0]) * sb
            v = (yb[:, k] - cb[1]) * sb
            A[:, k] = np.stack([u * x, u * y, u, v * x, v * y, v, x, y, np.ones(m)], axis=1)
        perm = np.tile(np.arange(9), (m, 1))
        for c in range(8):
            blk = np.abs(A[:, c:, c:]).reshape(m, -1)
            flat = np.argmax(blk, axis=1)
            best = blk[r_, flat]
            valid &= best >= 1e-12
            pr = c + flat // (9 - c)
            pc = c + flat % (9 - c)
            rc, rp = A[r_, c].copy(), A[r_, pr].copy()
            A[r_, c], A[r_, pr] = rp, rc
            cc, cp = A[r_, :, c].copy(), A[r_, :, pc].copy()
            A[r_, :, c], A[r_, :, pc] = cp, cc
            qc, qp = perm[r_, c].copy(), perm[r_, pc].copy()
            perm[r_, c], perm[r_, pc] = qp, qc
            for r in range(c + 1, 8):
                f = A[:, r, c] / A[:, c, c]
                for k in range(c, 9):
                    A[:, r, k] = A[:, r, k] - f * A[:, c, k]
        fz = np.zeros((m, 9))
        fz[:, 8] = 1.0
        for c in range(7, -1, -1):
            s_ = np.zeros(m)
            for k in range(c + 1, 9):
                s_ = s_ + A[:, c, k] * fz[:, k]
            fz[:, c] = -s_ / A[:, c, c]
        fh = np.zeros((m, 9))
        fh[r_[:, None], perm] = fz
        fh = rank2_project(fh)
        z = np.zeros(m)
        o = np.ones(m)
        Ta = [[sa, z, -sa * ca[0]], [z, sa, -sa * ca[1]], [z, z, o]]
        Tb = [[sb, z, -sb * cb[0]], [z, sb, -sb * cb[1]], [z, z, o]]
        G = [[(fh[:, 3 * i] * Ta[0][j] + fh[:, 3 * i + 1] * Ta[1][j]) + fh[:, 3 * i + 2] * Ta[2][j] for j in range(3)]
             for i in range(3)]
        F = np.zeros((m, 9))
        for i in range(3):
            for j in range(3):
                F[:, 3 * i + j] = (Tb[0][i] * G[0][j] + Tb[1][i] * G[1][j]) + Tb[2][i] * G[2][j]
    return F, valid


def inliers_f(F, a, b, thr2):
    """F (m, 9); a, b (n, 2) -> (m, n) bool: max squared point-to-epipolar-line distance < thr2,
    tested as d^2 < thr2 * |line normal|^2 on both lines (no division), with verify.hip's fused
    multiply-adds: lines (a2, b2, c2) = F x_a and (a1, b1) = the first two of F^T x_b, each
    entry fma(f_i, x, fma(f_j, y, f_k)); normals n = fma(a, a, b * b); the one residual
    d = x_b^T F x_a = fma(u, a2, fma(v, b2, c2)) for both lines: d^2 < thr2 * fmin(n1, n2)."""
    x, y, u, v = a[None, :, 0], a[None, :, 1], b[None, :, 0], b[None, :, 1]
    f = [F[:, k:k + 1] for k in range(9)]
    with np.errstate(divide="ignore", invalid="ignore", over="ignore"):
        a2, b2, c2 = fma(f[0], x, fma(f[1], y, f[2])), fma(f[3], x, fma(f[4], y, f[5])), fma(f[6], x, fma(f[7], y, f[8]))
        a1, b1 = fma(f[0], u, fma(f[3], v, f[6])), fma(f[1], u, fma(f[4], v, f[7]))
        n2 = fma(a2, a2, b2 * b2)
        n1 = fma(a1, a1, b1 * b1)
        d = fma(u, a2, fma(v, b2, c2))
        return d * d < thr2 * np.fmin(n1, n2)


def solve_s2(xa, ya, xb, yb):
    """Batched 2-point similarity: arrays (m, 2) -> (M (m, 9) = a, -b, tx, b, a, ty, 0, 0, 1,
    valid)."""
    dx, dy = xa[:, 1] - xa[:, 0], ya[:, 1] - ya[:, 0]
    ex, ey = xb[:, 1] - xb[:, 0], yb[:, 1] - yb[:, 0]
    den = dx * dx + dy * dy
    valid = den > 1e-12
    with np.errstate(divide="ignore", invalid="ignore"):
        a = (dx * ex + dy * ey) / den
        b = (dx * ey - dy * ex) / den
    M = np.zeros((len(a), 9))
    M[:, 0], M[:, 1], M[:, 3], M[:, 4], M[:, 8] = a, -b, b, a, 1.0
    M[:, 2] = xb[:, 0] - (a * xa[:, 0] - b * ya[:, 0])
    M[:, 5] = yb[:, 0] - (b * xa[:, 0] + a * ya[:, 0])
    return M, valid


def inliers_s(M, a, b, thr2):
    """M (m, 9); a, b (n, 2) -> (m, n) bool: |M x_a - x_b|^2 < thr2, verify.hip's fused form:
    dx = fma(m0, x, fma(m1, y, m2)) - u, dy likewise, fma(dx, dx, dy * dy)."""
    x, y, u, v = a[None, :, 0], a[None, :, 1], b[None, :, 0], b[None, :, 1]
    m = [M[:, k:k + 1] for k in range(6)]
    with np.errstate(invalid="ignore", over="ignore"):
        dx = fma(m[0], x, fma(m[1], y, m[2])) - u
        dy = fma(m[3], x, fma(m[4], y, m[5])) - v
        return fma(dx, dx, dy * dy) < thr2


def _lane_sums(vals, flags, lanes=64):
    """verify.hip lane_sums over matches: vals (n, E) with non-inliers as 0.0; lane l adds rows
    l, l + 64, ... in order, then the 64 lane sums are added in lane order -> (E,)."""
    n, E = vals.shape
    v = np.where(flags[:, None], vals, 0.0)
    R = -(-n // lanes)
    pad = np.zeros((R * lanes, E))
    pad[:n] = v
    pad = pad.reshape(R, lanes, E)
    part = np.zeros((lanes, E))
    for r in range(R):
        part = part + pad[r]
    tot = np.zeros(E)
    for l in range(lanes):
        tot = tot + part[l]
    return tot


def refit_h(a, b, flags, h):
    """verify.hip k_refit_h: the homography h (9,) refit on the inliers `flags` by normalized
    least squares (h33 = 1), or h unchanged (fewer than 5 inliers or a degenerate system)."""
    cnt = int(flags.sum())
    if cnt < 5:
        return h
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    dn = float(cnt)
    c4 = _lane_sums(np.c_[a, b], flags)
    cax, cay, cbx, cby = c4[0] / dn, c4[1] / dn, c4[2] / dn, c4[3] / dn
    dxa, dya = a[:, 0] - cax, a[:, 1] - cay
    dxb, dyb = b[:, 0] - cbx, b[:, 1] - cby
    r2 = _lane_sums(np.c_[np.sqrt(dxa * dxa + dya * dya), np.sqrt(dxb * dxb + dyb * dyb)], flags)
    ra, rb = r2[0], r2[1]
    if not (ra > 1e-12 and rb > 1e-12):
        return h
    sa = 1.4142135623730951 * dn / ra
    sb = 1.4142135623730951 * dn / rb
    x, y = (a[:, 0] - cax) * sa, (a[:, 1] - cay) * sa
    u, v = (b[:, 0] - cbx) * sb, (b[:, 1] - cby) * sb
    z, o = np.zeros(len(a)), np.ones(len(a))
    r0 = [x, y, o, z, z, z, -(u * x), -(u * y)]
    r1 = [z, z, z, x, y, o, -(v * x), -(v * y)]
    cols = [r0[i] * r0[j] + r1[i] * r1[j] for i in range(8) for j in range(i, 8)]
    cols += [r0[i] * u + r1[i] * v for i in range(8)]
    tot = _lane_sums(np.stack(cols, axis=1), flags)
    N = np.zeros((8, 9))
    e = 0
    for i in range(8):
        for j in range(i, 8):
            N[i, j] = N[j, i] = tot[e]
            e += 1
    N[:, 8] = tot[36:]
    tol = 1e-12 * max(0.0, N[np.arange(8), np.arange(8)].max())
    for c in range(8):
        col = np.abs(N[c:, c])
        piv = c + int(np.argmax(col))
        if not (col[piv - c] > tol):
            return h
        if piv != c:
            N[[c, piv]] = N[[piv, c]]
        for r in range(c + 1, 8):
            f = N[r, c] / N[c, c]
            N[r, c:] = N[r, c:] - f * N[c, c:]
    hh = np.zeros(9)
    for c in range(7, -1, -1):
        s_ = N[c, 8]
        for k in range(c + 1, 8):
            s_ = s_ - N[c, k] * hh[k]
        hh[c] = s_ / N[c, c]
    hh[8] = 1.0
    Ta = [[sa, 0.0, -sa * cax], [0.0, sa, -sa * cay], [0.0, 0.0, 1.0]]
    ib = 1.0 / sb
    Tbi = [[ib, 0.0, cbx], [0.0, ib, cby], [0.0, 0.0, 1.0]]
    G = [[(hh[3 * i] * Ta[0][j] + hh[3 * i + 1] * Ta[1][j]) + hh[3 * i + 2] * Ta[2][j] for j in range(3)]
         for i in range(3)]
    Hr = np.array([(Tbi[i][0] * G[0][j] + Tbi[i][1] * G[1][j]) + Tbi[i][2] * G[2][j] for i in range(3)
                   for j in range(3)])
    hmax = np.abs(Hr).max()
    if not (abs(Hr[8]) > 1e-12 * hmax):
        return h
    with np.errstate(all="ignore"):
        Hr = Hr / Hr[8]
    if not np.isfinite(Hr).all():
        return h
    return Hr


MODELS = {"homography": (4, solve_h4, inliers), "fundamental": (8, solve_f8, inliers_f),
          "similarity": (2, solve_s2, inliers_s)}


def ransac_pair(p, a, b, n_hyp, thr, seed, model="homography"):
    """Returns (inlier flags (n,), model (9,), count) for pair index p."""
    K, solve, test = MODELS[model]
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = len(a)
    if n < K:
        return np.zeros(n, bool), np.zeros(9), 0
    hyps = np.arange(n_hyp)
    idx, ok = samples(seed, p, n, hyps, K)
    safe = np.where(ok[:, None], idx, 0)
    h, valid = solve(a[safe, 0], a[safe, 1], b[safe, 0], b[safe, 1])
    valid &= ok
    cnt = np.where(valid, test(h, a, b, thr * thr).sum(1), -1)
    if cnt.max() < 0:
        return np.zeros(n, bool), np.zeros(9), 0
    best = int(np.argmax(cnt))          # first maximum = lowest hypothesis index
    flags = test(h[best:best + 1], a, b, thr * thr)[0]
    model_out = refit_h(a, b, flags, h[best]) if model == "homography" else h[best]
    return flags, model_out, int(cnt[best])


def confident(c, n, K, rounds, omc):
    """The adaptive stopping rule of fsx_ransac_batch_adaptive (verify.hip ransac_confident):
    (1 - w^K)^(256 rounds) <= omc (= 1 - confidence) with w = c / n, by the same float64
    multiplications in the same order: w^K by K - 1 successive products, the 256th power by 8
    squarings, then rounds - 1 products."""
    w = float(c) / float(n)
    wk = w
    for _ in range(K - 1):
        wk = wk * w
    q = 1.0 - wk
    for _ in range(8):
        q = q * q
    y = q
    for _ in range(rounds - 1):
        y = y * q
    return y <= omc


def ransac_pair_adaptive(p, a, b, max_hyp, conf, thr, seed, model="homography"):
    """The adaptive form (fsx_ransac_batch_adaptive) for pair index p: hypotheses in rounds of
    256 (at most max_hyp), the best so far kept (most inliers, then the lowest hypothesis);
    after each round the pair stops when it has drawn max_hyp, has fewer than K matches, or
    its best count c > 0 satisfies confident(c, n, K, rounds done, 1 - conf).  Returns
    (inlier flags (n,), model (9,), count, hypotheses drawn)."""
    K, solve, test = MODELS[model]
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    n = len(a)
    omc = 1.0 - conf
    best_c, best_h = -1, None
    used = 0
    for r in range((max_hyp + 255) // 256):
        h1 = min(256 * (r + 1), max_hyp)
        used = h1
        if n >= K:
            idx, ok = samples(seed, p, n, np.arange(256 * r, h1), K)
            safe = np.where(ok[:, None], idx, 0)
            h, valid = solve(a[safe, 0], a[safe, 1], b[safe, 0], b[safe, 1])
            valid &= ok
            cnt = np.where(valid, test(h, a, b, thr * thr).sum(1), -1)
            k = int(np.argmax(cnt))             # first maximum = lowest hypothesis of the round
            if cnt[k] > best_c:                 # equal counts keep the earlier round's winner
                best_c, best_h = int(cnt[k]), h[k]
        if h1 >= max_hyp or n < K or (best_c > 0 and confident(best_c, n, K, h1 // 256, omc)):
            break
    if best_c < 0:
        return np.zeros(n, bool), np.zeros(9), 0, used
    flags = test(best_h[None], a, b, thr * thr)[0]
    model_out = refit_h(a, b, flags, best_h) if model == "homography" else best_h
    return flags, model_out, best_c, used
