"""Write the golden vectors of the SIFT front end from the reference's own CPU path.

Run in the build container only (it imports ``planesite.gsift`` from the read-only reference
checkout, whose CPU path needs just numpy and scipy):

    python tests/golden/make_golden.py /root/reference/app

It writes small compressed ``.npz`` fixtures next to this script.  Inputs are the seeded
synthetic images of ``fieldstitch_amd.synth`` (identified by their SHA-256 so that a change of
the generator is caught), outputs are the reference's keypoint dictionaries, descriptors and
per-stage intermediates.  Nothing from the reference is stored except these numbers.
"""
import sys

sys.dont_write_bytecode = True      # nothing may be written into the reference checkout

import hashlib  # noqa: E402
import os  # noqa: E402
import time  # noqa: E402

import numpy as np  # noqa: E402
import scipy  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from fieldstitch_amd import synth  # noqa: E402


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def stage_intermediates(gsift, img, contrast, edge, n_hist=64):
    """Per-stage outputs of the reference's CPU path for one small image."""
    import scipy.ndimage as ndi
    gpyr, dpyr = gsift._pyramid(np, ndi, img)
    A = gsift._Atlas(np, gpyr, dpyr)
    F = gsift._extrema(np, ndi, dpyr, contrast, A)
    o, l, r, c, X, resp = gsift._refine(np, A, F, contrast, edge)
    scl = gsift.SIGMA * 2.0 ** ((l.astype(np.float32) + X[:, 2]) / gsift.N_LAYERS)
    # 36-bin histograms of the first n_hist refined keypoints (same math as _orientations_flat)
    R = gsift.ORI_R_MAX
    off = np.arange(-R, R + 1, dtype=np.int64)
    oi, oj = np.meshgrid(off, off, indexing="ij")
    oi = oi.ravel()[None]
    oj = oj.ravel()[None]
    d2 = (oi * oi + oj * oj).astype(np.float32)
    sl = slice(0, n_hist)
    oo, ll, rr, cc, ss = o[sl], l[sl], r[sl], c[sl], scl[sl]
    rows = A.rows[oo][:, None]
    cols = A.cols[oo][:, None]
    rad = np.rint(gsift.ORI_RADIUS * ss).astype(np.int64)[:, None]
    y = rr[:, None] + oi
    x = cc[:, None] + oj
    valid = (np.abs(oi) <= rad) & (np.abs(oj) <= rad) & (y > 0) & (y < rows - 1) & (x > 0) & (x < cols - 1)
    y = np.clip(y, 0, rows - 1)
    x = np.clip(x, 0, cols - 1)
    gi = A.gbase[oo][:, None] + (ll[:, None] - 1) * rows * cols + y * cols + x
    sw = gsift.ORI_SIG * ss
    w = np.exp(-d2 / (2.0 * sw[:, None] ** 2))
    m = A.MAG[gi] * w * valid
    b = np.rint(A.ANG[gi] * (gsift.ORI_BINS / 360.0)).astype(np.int64) % gsift.ORI_BINS
    n = rr.size
    idx = (np.arange(n, dtype=np.int64)[:, None] * gsift.ORI_BINS + b).ravel()
    hist = np.bincount(idx, weights=m.ravel(), minlength=n * gsift.ORI_BINS).astype(np.float32).reshape(n, -1)
    dog_sums = np.array([float(np.sum(D.astype(np.float64))) for D in dpyr])
    g_sums = np.array([[float(np.sum(G.astype(np.float64))) for G in layers] for layers in gpyr])
    return {
        "F": F.astype(np.int64), "ref_o": o.astype(np.int32), "ref_l": l.astype(np.int32),
        "ref_r": r.astype(np.int32), "ref_c": c.astype(np.int32), "ref_X": X.astype(np.float32),
        "ref_resp": resp.astype(np.float32), "scl": scl.astype(np.float32), "hist": hist,
        "dog_sums": dog_sums, "g_sums": g_sums,
        "dog_oct0": np.stack([D for D in dpyr[0]]) if img.size <= 100 * 100 else np.zeros(0, np.float32),
        "rows": A.rows_h.astype(np.int32), "cols": A.cols_h.astype(np.int32),
        "mag_first": A.MAG[:4096].astype(np.float32), "ang_first": A.ANG[:4096].astype(np.float32),
    }


def main(ref_app):
    sys.path.insert(0, ref_app)
    from planesite import gsift  # noqa: E402  (CPU path: numpy + scipy only)

    meta = {"numpy": np.__version__, "scipy": scipy.__version__}
    img1, img2, Hm = synth.field_pair(1)

    cases = [
        # name, image, crop, contrast, edge, max_kp, n_desc (None = all)
        ("c1_crop_cv", img1[300:540, 500:820], 0.04, 10.0, 12000, None),
        ("c1_crop_pipe", img1[300:540, 500:820], 0.015, 12.0, 12000, 384),
        ("c1_full_cv", img1, 0.04, 10.0, 4000, 384),
        ("c1_full2_cv", img2, 0.04, 10.0, 4000, 384),
        ("tiny_cv", img1[100:164, 200:280], 0.04, 10.0, 12000, None),
    ]
    for name, img, contrast, edge, max_kp, n_desc in cases:
        img = np.ascontiguousarray(img)
        t0 = time.time()
        kps, state = gsift.detect(img, gpu=False, max_kp=max_kp, contrast=contrast, edge=edge)
        n = len(kps["x"])
        sel = np.arange(n if n_desc is None else min(n, n_desc), dtype=np.int64)
        desc = gsift.describe(state, kps, sel)
        out = {"img_sha": np.array(sha(img)), "shape": np.array(img.shape, np.int32),
               "contrast": np.float32(contrast), "edge": np.float32(edge), "max_kp": np.int32(max_kp),
               "sel": sel, "desc": desc.astype(np.float32)}
        for k, v in kps.items():
            out["kp_" + k] = v
        if img.size <= 400 * 300:
            out["img"] = img
            for k, v in stage_intermediates(gsift, img, contrast, edge).items():
                out["st_" + k] = v
        out.update({"meta_" + k: np.array(v) for k, v in meta.items()})
        path = os.path.join(HERE, "gsift_%s.npz" % name)
        np.savez_compressed(path, **out)
        print("%-14s %4dx%-4d kps=%5d desc=%4d  %.1fs -> %s (%d KB)" % (
            name, img.shape[0], img.shape[1], n, len(sel), time.time() - t0, os.path.basename(path),
            os.path.getsize(path) // 1024))


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference/app")
