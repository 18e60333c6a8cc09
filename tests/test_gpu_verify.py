"""GPU homography RANSAC equals its CPU restatement exactly; filter feature rows."""
import numpy as np
import pytest

from oracle import ransac_oracle as ro

pytestmark = pytest.mark.gpu


def _planted(rng, n, frac, H):
    a = np.c_[rng.uniform(0, 2640, n), rng.uniform(0, 1978, n)]
    w = H[2, 0] * a[:, 0] + H[2, 1] * a[:, 1] + H[2, 2]
    b = np.c_[(H[0, 0] * a[:, 0] + H[0, 1] * a[:, 1] + H[0, 2]) / w, (H[1, 0] * a[:, 0] + H[1, 1] * a[:, 1] + H[1, 2]) / w]
    b += rng.normal(0, 0.7, b.shape)
    bad = rng.random(n) >= frac
    b[bad] = np.c_[rng.uniform(0, 2640, bad.sum()), rng.uniform(0, 1978, bad.sum())]
    return a, b


def test_ransac_equals_oracle(gpu):
    from fieldstitch_amd.verify import ransac_homography
    rng = np.random.default_rng(4)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    sets = [_planted(rng, n, f, H) for n, f in ((900, 0.6), (300, 0.3), (3, 1.0), (0, 1.0), (2000, 0.9))]
    line = np.c_[np.arange(30.0), 2 * np.arange(30.0)]
    sets.append((line, line + 5.0))
    flags, Hs, cnt = ransac_homography([a for a, _ in sets], [b for _, b in sets], n_hyp=384, thr=3.0, seed=99)
    for p, (a, b) in enumerate(sets):
        rf, rh, rc = ro.ransac_pair(p, a, b, 384, 3.0, 99)
        print("pair", p, len(a), rc, flush=True)
        assert cnt[p] == rc
        np.testing.assert_array_equal(flags[p], rf)
        np.testing.assert_allclose(Hs[p].reshape(9), rh, rtol=1e-12, atol=1e-12)


def _views(rng, n, t, frac, noise=0.3):
    f, c = 1500.0, (1300.0, 980.0)
    K = np.array([[f, 0.0, c[0]], [0.0, f, c[1]], [0.0, 0.0, 1.0]])
    az = np.radians(3.0)
    R = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    X = np.c_[rng.uniform(-60, 60, n), rng.uniform(-45, 45, n), 100.0 + rng.uniform(-8, 8, n)]
    xa = (K @ X.T).T
    xb = (K @ (R @ (X - np.asarray(t, float)).T)).T
    a, b = xa[:, :2] / xa[:, 2:], xb[:, :2] / xb[:, 2:]
    b = b + rng.normal(0.0, noise, b.shape)
    bad = rng.random(n) >= frac
    b[bad] = rng.uniform(0, 2600, (int(bad.sum()), 2))
    return a, b


@pytest.mark.parametrize("model,thr", [("fundamental", 1.0), ("similarity", 4.0)])
def test_ransac_models_equal_oracle(gpu, model, thr):
    from fieldstitch_amd.verify import ransac
    rng = np.random.default_rng(12)
    sets = [_views(rng, 700, (12.0, 3.0, 1.0), 0.7), _views(rng, 250, (20.0, 2.0, 0.0), 0.5),
            _views(rng, 9, (5.0, 1.0, 0.0), 1.0), (np.zeros((0, 2)), np.zeros((0, 2))),
            _views(rng, 1500, (2.0, 9.0, 0.3), 0.9)]
    pt = np.tile([[100.0, 200.0]], (20, 1))
    sets.append((pt, pt + 5.0))
    flags, Ms, cnt = ransac(model, [a for a, _ in sets], [b for _, b in sets], n_hyp=320, thr=thr, seed=1234)
    for p, (a, b) in enumerate(sets):
        rf, rm, rc = ro.ransac_pair(p, a, b, 320, thr, 1234, model=model)
        print("pair", model, p, len(a), rc, flush=True)
        assert cnt[p] == rc
        np.testing.assert_array_equal(flags[p], rf)
        np.testing.assert_allclose(Ms[p], rm, rtol=1e-10, atol=1e-12)
    if model == "fundamental":
        assert cnt[0] > 0.5 * 700 and cnt[4] > 0.85 * 1500


def test_gather_and_similarity_fit(gpu):
    """fsx_gather_xy against numpy indexing (out-of-range indices give NaN rows), and
    fsx_similarity_fit_batch against numpy's least-squares solve of the same similarity on
    each pair's inliers (fewer than 2 inliers and coincident points give NaN)."""
    import ctypes as C
    import math
    import torch
    from fieldstitch_amd import _lib
    rng = np.random.default_rng(21)
    xy = rng.uniform(-500, 3000, (1000, 2))
    idx = rng.integers(0, 1000, 5000).astype(np.int32)
    idx[[3, 77, 4999]] = [-1, 1000, 2 ** 31 - 1]
    d_xy = torch.from_numpy(xy).cuda()
    d_idx = torch.from_numpy(idx).cuda()
    out = torch.empty((5000, 2), dtype=torch.float64, device="cuda")
    P = lambda t: C.c_void_p(t.data_ptr())    # noqa: E731
    st = C.c_void_p(_lib.stream_handle())
    _lib.call("fsx_gather_xy", 5000, P(d_idx), P(d_xy), 1000, P(out), st)
    g = out.cpu().numpy()
    bad = (idx < 0) | (idx >= 1000)
    np.testing.assert_array_equal(g[~bad], xy[idx[~bad]])
    assert np.isnan(g[bad]).all()

    sets, flags = [], []
    for k, n in enumerate((300, 2, 1, 0, 50, 40)):
        a = rng.uniform(0, 2000, (n, 2))
        ang, sc = rng.uniform(-0.5, 0.5), rng.uniform(0.8, 1.2)
        R = sc * np.array([[math.cos(ang), -math.sin(ang)], [math.sin(ang), math.cos(ang)]])
        b = a @ R.T + rng.uniform(-100, 100, 2) + rng.normal(0, 0.5, (n, 2))
        f = rng.random(n) < 0.8
        if k == 1:
            f[:] = True
        if k == 5:
            a[:] = a[0]                     # coincident points
        sets.append((a, b))
        flags.append(f)
    off = np.r_[0, np.cumsum([len(a) for a, _ in sets])].astype(np.int64)
    A = np.concatenate([a for a, _ in sets]) if off[-1] else np.zeros((1, 2))
    B = np.concatenate([b for _, b in sets])
    Fl = np.concatenate(flags).astype(np.uint8)
    d = [torch.from_numpy(np.ascontiguousarray(x)).cuda() for x in (off, A, B, Fl)]
    fit = torch.empty((len(sets), 4), dtype=torch.float64, device="cuda")
    _lib.call("fsx_similarity_fit_batch", len(sets), P(d[0]), P(d[1]), P(d[2]), P(d[3]), P(fit), st)
    fit = fit.cpu().numpy()
    for k, ((a, b), f) in enumerate(zip(sets, flags)):
        a, b = a[f], b[f]
        if len(a) < 2 or k == 5:
            assert np.isnan(fit[k]).all(), k
            continue
        M = np.zeros((2 * len(a), 4))
        M[0::2] = np.c_[a[:, 0], -a[:, 1], np.ones(len(a)), np.zeros(len(a))]
        M[1::2] = np.c_[a[:, 1], a[:, 0], np.zeros(len(a)), np.ones(len(a))]
        (sa, sb, tx, ty), *_ = np.linalg.lstsq(M, b.reshape(-1), rcond=None)
        want = (math.atan2(sb, sa), math.hypot(sa, sb), tx, ty)
        np.testing.assert_allclose(fit[k], want, rtol=1e-9, atol=1e-7)


@pytest.mark.parametrize("model,thr", [("homography", 3.0), ("fundamental", 1.0), ("similarity", 4.0)])
def test_small_batch_split_equals_one_block_per_pair(gpu, model, thr):
    """A small batch scores each pair's hypotheses on several blocks and picks the winner in a
    second launch; a batch of 600+ pairs uses one block per pair.  The first pairs of both
    (same slots, same draws) give identical flags, counts and models."""
    from fieldstitch_amd.verify import ransac
    rng = np.random.default_rng(21)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    if model == "fundamental":
        sets = [_views(rng, n, (12.0, 3.0, 1.0), f) for n, f in ((700, 0.7), (60, 0.5), (1500, 0.9))]
    else:
        sets = [_planted(rng, n, f, H if model == "homography" else np.array([[0.9, -0.2, 30.0], [0.2, 0.9, 5.0],
                                                                             [0.0, 0.0, 1.0]]))
                for n, f in ((700, 0.7), (60, 0.5), (1500, 0.9))]
    pad = [sets[0]] * 620
    print("case split", model, flush=True)
    small = ransac(model, [a for a, _ in sets], [b for _, b in sets], n_hyp=2000, thr=thr, seed=77)
    big = ransac(model, [a for a, _ in sets + pad], [b for _, b in sets + pad], n_hyp=2000, thr=thr, seed=77)
    for p in range(len(sets)):
        np.testing.assert_array_equal(small[0][p], big[0][p])
        np.testing.assert_array_equal(small[1][p], big[1][p])
        assert small[2][p] == big[2][p] and small[2][p] > 0


@pytest.mark.parametrize("model,thr,conf,cap", [("homography", 3.0, 0.995, 2000), ("fundamental", 1.0, 0.999, 1500),
                                                 ("similarity", 4.0, 0.99, 2000)])
def test_adaptive_batch_equals_oracle(gpu, model, thr, conf, cap):
    """fsx_ransac_batch_adaptive against ransac_oracle.ransac_pair_adaptive, pair by pair:
    counts, inlier flags, models and hypotheses drawn, on sets that stop after one round (all
    inliers), after a few (70 %), that never stop before the cap (few inliers), and sets with
    fewer matches than the sample size, none, or only degenerate samples."""
    import ctypes as C
    import torch
    from fieldstitch_amd import _lib
    from fieldstitch_amd.match_flow import _MODEL
    rng = np.random.default_rng(41)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    if model == "homography":
        sets = [_planted(rng, 600, 1.0, H), _planted(rng, 900, 0.7, H), _planted(rng, 400, 0.25, H)]
    else:
        sets = [_views(rng, 600, (12.0, 3.0, 1.0), 1.0, noise=0.05), _views(rng, 900, (20.0, 2.0, 0.0), 0.7),
                _views(rng, 400, (2.0, 9.0, 0.3), 0.25)]
    sets += [(sets[1][0][:3], sets[1][1][:3]), (np.zeros((0, 2)), np.zeros((0, 2)))]
    pt = np.tile([[100.0, 200.0]], (20, 1))
    sets.append((pt, pt + 5.0))
    n = len(sets)
    off = np.zeros(n + 1, np.int64)
    off[1:] = np.cumsum([len(a) for a, _ in sets])
    A = torch.from_numpy(np.concatenate([a for a, _ in sets] + [np.zeros((1, 2))])).cuda()
    B = torch.from_numpy(np.concatenate([b for _, b in sets] + [np.zeros((1, 2))])).cuda()
    seeds = [0x1234 + 977 * k for k in range(n)]
    d_off = torch.from_numpy(off).cuda()
    d_thr = torch.full((n,), thr, dtype=torch.float64, device="cuda")
    d_seed = torch.tensor(seeds, dtype=torch.int64, device="cuda")
    inl = torch.zeros(int(off[-1]) + 1, dtype=torch.uint8, device="cuda")
    M = torch.zeros((n, 9), dtype=torch.float64, device="cuda")
    cnt = torch.zeros(n, dtype=torch.int32, device="cuda")
    used = torch.full((n,), -1, dtype=torch.int32, device="cuda")
    P = C.c_void_p
    print("adaptive", model, [len(a) for a, _ in sets], flush=True)
    _lib.call("fsx_ransac_batch_adaptive", _MODEL[model], n, P(d_off.data_ptr()), P(A.data_ptr()), P(B.data_ptr()),
              cap, conf, P(d_thr.data_ptr()), P(d_seed.data_ptr()), P(inl.data_ptr()), P(M.data_ptr()),
              P(cnt.data_ptr()), P(used.data_ptr()), P(_lib.stream_handle()))
    torch.cuda.synchronize()
    flags, Ms, cs, us = inl.cpu().numpy().view(bool), M.cpu().numpy(), cnt.cpu().numpy(), used.cpu().numpy()
    rounds = []
    for p, (a, b) in enumerate(sets):
        rf, rm, rc, ru = ro.ransac_pair_adaptive(0, a, b, cap, conf, thr, seeds[p], model=model)
        print("pair", p, len(a), rc, ru, flush=True)
        assert (cs[p], us[p]) == (rc, ru)
        np.testing.assert_array_equal(flags[off[p]:off[p + 1]], rf)
        np.testing.assert_allclose(Ms[p], rm, rtol=1e-10, atol=1e-12)
        rounds.append(ru)
    # one round with all inliers; a degenerate set never gets a model and runs to the cap; few
    # inliers reach the cap for the 4- and 8-point models (the 2-point one stops after 512)
    assert rounds[0] == 256 and 256 <= rounds[1] < cap and rounds[-1] == cap
    assert rounds[2] == (512 if model == "similarity" else cap)
