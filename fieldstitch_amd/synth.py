"""Seeded synthetic inputs for the SIFT front end and the guided matcher.

Every image is a procedural "field" mosaic evaluated at the mosaic coordinates of each image
pixel: fBm value noise (6 octaves, persistence 0.5) plus random filled polygons with a
uniform brightness offset, plus fresh Gaussian noise per image, clipped to uint8.  Because
the mosaic is a function rather than a stored raster, a crop, a rotated/scaled crop and a
homography-warped copy are all exact (no resampling blur), and the true image-to-image
homography of every pair is known.

Workloads (sizes are the ones the benchmarks and parity tests use):

* ``field_pair``        one 1368x912 image plus its copy warped by a known homography
                         (rotation 5 deg, scale 1.03, x-translation 30 % of the width);
* ``flight``            photos cropped along a flight grid (front/side overlap, rotation
                         and scale jitter), with per-pair true homographies and a GNSS-style
                         prior (true mapping + N(0, sigma) px offset per pair);
* ``sift_like_descriptors``  SIFT-like float descriptor pairs with planted true matches.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

_M32 = np.uint64(0xFFFFFFFF)


def _hash2(ix: np.ndarray, iy: np.ndarray, seed: int) -> np.ndarray:
    """Integer lattice hash -> float32 in [0, 1). ix, iy int64 arrays (any sign)."""
    h = (ix.astype(np.uint64) * np.uint64(0x8DA6B343)) ^ (iy.astype(np.uint64) * np.uint64(0xD8163841))
    h = (h ^ np.uint64((seed * 0xCB1AB31F) & 0xFFFFFFFF)) & _M32
    h ^= h >> np.uint64(16)
    h = (h * np.uint64(0x85EBCA6B)) & _M32
    h ^= h >> np.uint64(13)
    h = (h * np.uint64(0xC2B2AE35)) & _M32
    h ^= h >> np.uint64(16)
    return (h.astype(np.float64) * (1.0 / 4294967296.0)).astype(np.float32)


def _value_noise(u: np.ndarray, v: np.ndarray, period: float, seed: int) -> np.ndarray:
    """Smoothstep-interpolated value noise on a lattice of the given period (mosaic px)."""
    fu = u / period
    fv = v / period
    i0 = np.floor(fu)
    j0 = np.floor(fv)
    tx = (fu - i0).astype(np.float32)
    ty = (fv - j0).astype(np.float32)
    i0 = i0.astype(np.int64)
    j0 = j0.astype(np.int64)
    # lattice values for the covering box only, then gather the 4 corners
    imin, imax = int(i0.min()), int(i0.max()) + 1
    jmin, jmax = int(j0.min()), int(j0.max()) + 1
    gi, gj = np.meshgrid(np.arange(imin, imax + 1), np.arange(jmin, jmax + 1), indexing="xy")
    lat = _hash2(gi, gj, seed)                   # shape (nj, ni)
    a = i0 - imin
    b = j0 - jmin
    h00 = lat[b, a]
    h10 = lat[b, a + 1]
    h01 = lat[b + 1, a]
    h11 = lat[b + 1, a + 1]
    sx = tx * tx * (3.0 - 2.0 * tx)
    sy = ty * ty * (3.0 - 2.0 * ty)
    top = h00 + (h10 - h00) * sx
    bot = h01 + (h11 - h01) * sx
    return top + (bot - top) * sy


def fbm(u: np.ndarray, v: np.ndarray, seed: int, octaves: int = 6, persistence: float = 0.5,
        base_period: float = 20.0) -> np.ndarray:
    """fBm value noise in [0, 1] at mosaic coordinates (u, v)."""
    acc = np.zeros(u.shape, np.float32)
    amp, tot, period = 1.0, 0.0, base_period
    for k in range(octaves):
        acc += np.float32(amp) * _value_noise(u, v, period, seed * 131 + k)
        tot += amp
        amp *= persistence
        period *= 0.5
    return acc * np.float32(1.0 / tot)


@dataclass
class Polygon:
    verts: np.ndarray      # (k, 2) float64 mosaic (u, v)
    offset: float          # brightness offset


def random_polygons(rng: np.random.Generator, n: int, u0: float, v0: float, u1: float, v1: float,
                    rmin: float = 15.0, rmax: float = 140.0) -> list:
    """n star-shaped (simple) polygons with centres uniform in [u0,u1)x[v0,v1)."""
    polys = []
    for _ in range(n):
        cu = rng.uniform(u0, u1)
        cv = rng.uniform(v0, v1)
        k = int(rng.integers(4, 10))
        ang = np.sort(rng.uniform(0, 2 * math.pi, k))
        rad = rng.uniform(rmin, rmax) * rng.uniform(0.45, 1.0, k)
        verts = np.c_[cu + rad * np.cos(ang), cv + rad * np.sin(ang)]
        polys.append(Polygon(verts, float(rng.uniform(-40.0, 40.0))))
    return polys


def _inside(verts: np.ndarray, u: np.ndarray, v: np.ndarray) -> np.ndarray:
    """Even-odd point-in-polygon test (crossing number)."""
    ins = np.zeros(u.shape, bool)
    k = len(verts)
    for a in range(k):
        xa, ya = verts[a]
        xb, yb = verts[(a + 1) % k]
        cond = (ya > v) != (yb > v)
        with np.errstate(divide="ignore", invalid="ignore"):
            xc = xa + (v - ya) * (xb - xa) / (yb - ya)
        ins ^= cond & (u < xc)
    return ins


@dataclass
class Mosaic:
    """A procedural field: fBm + polygons. Pixel values before noise are in [0, 255]."""
    seed: int
    polys: list = field(default_factory=list)
    lo: float = 0.0
    span: float = 255.0

    def render(self, A: np.ndarray, H: int, W: int, noise_rng: np.random.Generator, sigma: float = 2.0,
               gsd: float = 1.0) -> np.ndarray:
        """Render an H x W uint8 image whose pixel (x, y) shows mosaic point A @ [x, y, 1]
        (A: 2x3 affine, or 3x3 projective mapping image px -> mosaic px)."""
        yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
        P = np.asarray(A, np.float64)
        if P.shape == (3, 3):
            w = P[2, 0] * xx + P[2, 1] * yy + P[2, 2]
            u = (P[0, 0] * xx + P[0, 1] * yy + P[0, 2]) / w
            v = (P[1, 0] * xx + P[1, 1] * yy + P[1, 2]) / w
        else:
            u = P[0, 0] * xx + P[0, 1] * yy + P[0, 2]
            v = P[1, 0] * xx + P[1, 1] * yy + P[1, 2]
        img = np.float32(self.lo) + np.float32(self.span) * fbm(u, v, self.seed)
        Pf = P if P.shape == (3, 3) else np.vstack([P, [0.0, 0.0, 1.0]])
        Pinv = np.linalg.inv(Pf)
        for pg in self.polys:
            # polygon vertices in image px -> bounding box of the pixels to test
            q = np.c_[pg.verts, np.ones(len(pg.verts))] @ Pinv.T
            q = q[:, :2] / q[:, 2:3]
            x0 = max(int(math.floor(q[:, 0].min())) - 1, 0)
            x1 = min(int(math.ceil(q[:, 0].max())) + 2, W)
            y0 = max(int(math.floor(q[:, 1].min())) - 1, 0)
            y1 = min(int(math.ceil(q[:, 1].max())) + 2, H)
            if x0 >= x1 or y0 >= y1:
                continue
            ins = _inside(pg.verts, u[y0:y1, x0:x1], v[y0:y1, x0:x1])
            img[y0:y1, x0:x1][ins] += np.float32(pg.offset)
        img += noise_rng.normal(0.0, sigma, img.shape).astype(np.float32)
        return np.clip(np.rint(img), 0, 255).astype(np.uint8)


def similarity(theta_deg: float, scale: float, tx: float, ty: float, cx: float = 0.0, cy: float = 0.0) -> np.ndarray:
    """3x3 matrix p -> s R (p - c) + c + t."""
    t = math.radians(theta_deg)
    c, s = math.cos(t) * scale, math.sin(t) * scale
    M = np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])
    T1 = np.array([[1, 0, -cx], [0, 1, -cy], [0, 0, 1.0]])
    T2 = np.array([[1, 0, cx + tx], [0, 1, cy + ty], [0, 0, 1.0]])
    return T2 @ M @ T1


def field_pair(seed: int = 1, W: int = 1368, H: int = 912, n_poly: int = 40):
    """One field image and its copy warped by a known homography Hm (image-1 px -> image-2 px):
    rotation 5 deg and scale 1.03 about the centre, plus a translation of 30 % of the width.
    Returns (img1, img2, Hm)."""
    rng = np.random.default_rng(seed)
    mos = Mosaic(seed=int(rng.integers(1 << 30)))
    mos.polys = random_polygons(rng, n_poly, 0, 0, W, H)
    img1 = mos.render(np.eye(3)[:2], H, W, rng)
    Hm = similarity(5.0, 1.03, 0.3 * W, 0.0, (W - 1) / 2.0, (H - 1) / 2.0)
    img2 = mos.render(np.linalg.inv(Hm), H, W, rng)     # pixel of img2 -> img1 coords = mosaic coords
    return img1, img2, Hm


@dataclass
class Flight:
    images: list           # uint8 arrays
    A: list                # 3x3 image px -> mosaic px
    pairs: list            # (i, j) overlapping pairs, i < j
    H_true: dict           # (i, j) -> 3x3 true homography image i px -> image j px
    prior: dict            # (i, j) -> 3x3 prior mapping (true + N(0, sigma) translation)


def _render_job(job):
    mos, M, h, w, noise_seed = job
    return mos.render(M, h, w, np.random.default_rng(noise_seed))


def _pool_context():
    """Process start method of the render pools: fork, unless this process has initialised the
    GPU runtime (children forked from such a process can crash), then spawn (the workers import
    only numpy and this module)."""
    import multiprocessing as mp
    import sys
    torch = sys.modules.get("torch")
    if torch is not None and torch.cuda.is_initialized():
        return mp.get_context("spawn")
    return mp.get_context("fork")


def flight(seed: int = 2, rows: int = 2, cols: int = 4, W: int = 2640, H: int = 1978,
           front: float = 0.75, side: float = 0.65, rot_jitter: float = 3.0, scale_jitter: float = 0.02,
           prior_sigma: float = 3.0, lowres_frac: float = 0.0, poly_density: float = 40.0 / (1368 * 912),
           min_overlap: float = 0.05, workers: int = 1, progress=None, lowres_pattern=None) -> Flight:
    """Photos along a rows x cols flight grid.  Flight lines run along x (a line is one grid row),
    so consecutive photos of a line overlap by `front` of the width and neighbouring lines by
    `side` of the height.  A photo is low resolution (W/2 x H/2 covering the same ground) with
    probability `lowres_frac`.  Pairs: every photo pair whose footprints (without jitter)
    overlap by more than `min_overlap` of the footprint area.  Each photo's noise has its own
    seed, so `workers` > 1 renders in parallel with identical results.  progress(done, total)
    is called every 25 photos.  ``lowres_pattern`` (one bool per photo, row-major) fixes which
    photos are low resolution instead of the draw (the draw still happens, so the rest of the
    flight is the same)."""
    rng = np.random.default_rng(seed)
    step_u = (1.0 - front) * W
    step_v = (1.0 - side) * H
    ext_u = step_u * (cols - 1) + W
    ext_v = step_v * (rows - 1) + H
    margin = 0.1 * max(W, H)
    mos = Mosaic(seed=int(rng.integers(1 << 30)))
    n_poly = int(round(poly_density * (ext_u + 2 * margin) * (ext_v + 2 * margin)))
    mos.polys = random_polygons(rng, n_poly, -margin, -margin, ext_u + margin, ext_v + margin)
    jobs, A, centres = [], [], []
    for r in range(rows):
        for c in range(cols):
            theta = rng.uniform(-rot_jitter, rot_jitter)
            s = 1.0 + rng.uniform(-scale_jitter, scale_jitter)
            low = rng.random() < lowres_frac
            if lowres_pattern is not None:
                low = bool(lowres_pattern[r * cols + c])
            w, h, g = (W // 2, (H + 1) // 2, 2.0) if low else (W, H, 1.0)
            cu = c * step_u + W / 2.0
            cv = r * step_v + H / 2.0
            # image px (x, y) -> mosaic: (cu, cv) + s*g*R(theta) ((x, y) - image centre)
            M = similarity(theta, s * g, cu - (w - 1) / 2.0, cv - (h - 1) / 2.0, (w - 1) / 2.0, (h - 1) / 2.0)
            A.append(M)
            centres.append((cu, cv))
            jobs.append((mos, M, h, w, int(rng.integers(1 << 62))))
    if workers > 1:
        images = []
        with _pool_context().Pool(min(workers, len(jobs))) as pool:
            for im in pool.imap(_render_job, jobs):
                images.append(im)
                if progress and len(images) % 25 == 0:
                    progress(len(images), len(jobs))
    else:
        images = []
        for j in jobs:
            images.append(_render_job(j))
            if progress and len(images) % 25 == 0:
                progress(len(images), len(jobs))
    n = len(images)
    pairs, H_true, prior = [], {}, {}
    for i in range(n):
        for j in range(i + 1, n):
            du = abs(centres[i][0] - centres[j][0])
            dv = abs(centres[i][1] - centres[j][1])
            ov = max(0.0, W - du) * max(0.0, H - dv) / float(W * H)
            if ov > min_overlap:
                pairs.append((i, j))
                Hij = np.linalg.inv(A[j]) @ A[i]
                H_true[(i, j)] = Hij
                d = rng.normal(0.0, prior_sigma, 2)
                P = Hij.copy()
                P[0, 2] += d[0]
                P[1, 2] += d[1]
                prior[(i, j)] = P
    return Flight(images, A, pairs, H_true, prior)


def sift_like_descriptors(rng: np.random.Generator, n: int) -> np.ndarray:
    """SIFT-like float32 descriptors: 4x4x8 gamma(0.6, 1) histograms, L2-normalised,
    clipped at 0.2, renormalised, scaled by 512."""
    d = rng.gamma(0.6, 1.0, (n, 128))
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    d = np.minimum(d, 0.2)
    d /= np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-12)
    return (d * 512.0).astype(np.float32)


def descriptor_pair(rng: np.random.Generator, na: int, nb: int, n_true: int | None = None, noise: float = 8.0):
    """Two descriptor sets; the first n_true rows of a (shuffled into b) are perturbed copies
    (per-bin N(0, noise)), the rest independent.  Returns (da, db, truth) with truth = (qa, tb)."""
    n_true = min(na, nb) // 2 if n_true is None else n_true
    da = sift_like_descriptors(rng, na)
    db = sift_like_descriptors(rng, nb)
    tb = rng.permutation(nb)[:n_true]
    qa = rng.permutation(na)[:n_true]
    db[tb] = np.maximum(da[qa] + rng.normal(0.0, noise, (n_true, 128)).astype(np.float32), 0.0)
    return da, db, (qa, tb)


def flight_cameras(fl: Flight, f: float = 3000.0, gnss_sigma_px: float = 3.0 / np.sqrt(2.0), seed: int = 0):
    """Pinhole nadir cameras that see the flight's photos exactly (for the ground-plane prior
    of matching.py:155-172): the mosaic plane is the ground (z = 0, z pointing down), photo i
    looks straight down from height f * s_i (s_i = mosaic units per pixel) with its image axes
    along the photo's rotation, no lens distortion.  The camera centres carry GNSS noise of
    gnss_sigma_px pixels per axis (so a pair's prediction errs by ~gnss_sigma_px * sqrt(2)).
    Returns the matching context's camera entries: K, dist, R, C, zg, has_rot, W, H."""
    rng = np.random.default_rng(seed)
    K, dist, R, Cc, zg, W, H = [], [], [], [], [], [], []
    for A, img in zip(fl.A, fl.images):
        h, w = img.shape
        a, b = float(A[0, 0]), float(A[1, 0])
        s = float(np.hypot(a, b))
        c, sn = a / s, b / s
        cx, cy = (w - 1) / 2.0, (h - 1) / 2.0
        centre = np.asarray(A, np.float64)[:2] @ np.array([cx, cy, 1.0])
        K.append(np.array([[f, 0.0, cx], [0.0, f, cy], [0.0, 0.0, 1.0]]))
        dist.append(np.zeros(5))
        R.append(np.array([[c, sn, 0.0], [-sn, c, 0.0], [0.0, 0.0, 1.0]]))
        noise = rng.normal(0.0, gnss_sigma_px * s, 2)
        Cc.append(np.array([centre[0] + noise[0], centre[1] + noise[1], -f * s]))
        zg.append(0.0)
        W.append(w)
        H.append(h)
    n = len(fl.images)
    return {"K": K, "dist": dist, "R": R, "C": Cc, "zg": zg, "has_rot": [True] * n, "W": W, "H": H}
