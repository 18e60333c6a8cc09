"""The CPU restatement of match_pair's chain (oracle/match_pair_oracle.py) pinned on planted
scenes: every returned match is a true correspondence, most visible ones are found, the
similarity recovers the planted rotation / scale, and the reference's failure tuples come
back for too few features and for unrelated photos (CPU only)."""
import math

import numpy as np

from oracle import match_pair_oracle as mpo

W, H = 1320, 989
CTX = {"W": [W, W, W, W], "H": [H, H, H, H], "has_rot": [False] * 4, "ratio": 0.8, "min_inliers": 20,
       "ransac_thr": 1.0, "guided": True, "n_coarse": 1000, "guided_radius": 0.03, "guided_k": 16, "prior_tol": 0.18,
       "ransac_iters": 1024}


def _scene(seed, n=300, ang_deg=3.0, scale=1.02, shift=(260.0, 25.0), distractors=60):
    rng = np.random.default_rng(seed)
    xa = np.c_[rng.uniform(0, W, n), rng.uniform(0, H, n)]
    da = rng.integers(0, 61, (n, 128)).astype(np.uint8)
    a = math.radians(ang_deg)
    R = scale * np.array([[math.cos(a), -math.sin(a)], [math.sin(a), math.cos(a)]])
    xb_all = xa @ R.T + np.asarray(shift) + rng.normal(0, 0.3, (n, 2))
    vis = (xb_all[:, 0] >= 0) & (xb_all[:, 0] < W) & (xb_all[:, 1] >= 0) & (xb_all[:, 1] < H)
    src = np.nonzero(vis)[0]
    xb = xb_all[src]
    db = np.clip(da[src].astype(int) + rng.integers(-2, 3, (len(src), 128)), 0, 255).astype(np.uint8)
    xb = np.r_[xb, np.c_[rng.uniform(0, W, distractors), rng.uniform(0, H, distractors)]]
    db = np.r_[db, rng.integers(0, 61, (distractors, 128)).astype(np.uint8)]
    perm = rng.permutation(len(xb))
    truth = np.full(len(xb), -1)
    truth[:len(src)] = src
    return (da, xa), (db[perm], xb[perm]), truth[perm], len(src)


def test_chain_recovers_planted_correspondences():
    fa, fb, truth, n_vis = _scene(3)
    i, j, q, t, sim, n_in = mpo.match_pair(CTX, {0: fa, 1: fb}, 0, 1)
    assert (i, j) == (0, 1) and q is not None and q.dtype == np.int32 and t.dtype == np.int32
    assert n_in == len(q) == len(t)
    assert np.array_equal(truth[t], q)                 # every match is a true correspondence
    assert len(q) >= 0.9 * n_vis
    assert len(np.unique(t)) == len(t)                 # one query per target
    ang, sc, tx, ty = sim
    assert abs(ang - math.radians(3.0)) < 2e-3 and abs(sc - 1.02) < 2e-3
    assert abs(tx - 260.0) < 3.0 and abs(ty - 25.0) < 3.0


def test_chain_failure_tuples():
    fa, fb, _, _ = _scene(4)
    few = (fb[0][:20], fb[1][:20])
    assert mpo.match_pair(CTX, {0: fa, 1: few}, 0, 1) == (0, 1, None, None, None, 0)
    rng = np.random.default_rng(5)
    other = (rng.integers(0, 61, (300, 128)).astype(np.uint8), np.c_[rng.uniform(0, W, 300), rng.uniform(0, H, 300)])
    i, j, q, t, sim, n_in = mpo.match_pair(CTX, {0: fa, 1: other}, 0, 1)
    assert q is None and t is None and sim is None


def test_chain_without_guidance_uses_the_exhaustive_fallback():
    fa, fb, truth, n_vis = _scene(6, ang_deg=-2.0, scale=0.99, shift=(-200.0, 60.0))
    ctx = dict(CTX, guided=False)
    i, j, q, t, sim, n_in = mpo.match_pair(ctx, {0: fa, 1: fb}, 0, 1)
    assert q is not None and np.array_equal(truth[t], q) and len(q) >= 0.9 * n_vis
    assert abs(sim[0] - math.radians(-2.0)) < 2e-3 and abs(sim[1] - 0.99) < 2e-3
