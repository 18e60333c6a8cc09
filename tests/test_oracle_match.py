"""Pins of the matcher oracle (oracle/match_oracle.py): the FP64 oracle against a brute-force
loop and the float32 restatement of matching.py, the spatial candidate search against scipy's
cKDTree, and the tie rules on hand-made cases."""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import match_oracle as mo


def rootsift_like(rng, n):
    d = rng.gamma(0.6, 1.0, (n, 128))
    d /= d.sum(1, keepdims=True)
    return np.clip(np.sqrt(d) * 512, 0, 255).astype(np.uint8)


def pair_u8(rng, na, nb, n_true, noise=6):
    a = rootsift_like(rng, na)
    b = rootsift_like(rng, nb)
    qa = rng.permutation(na)[:n_true]
    tb = rng.permutation(nb)[:n_true]
    b[tb] = np.clip(a[qa].astype(int) + rng.integers(-noise, noise + 1, (n_true, 128)), 0, 255)
    return a, b


def brute_loop(da, db, ratio):
    da = np.asarray(da, np.float64)
    db = np.asarray(db, np.float64)
    best = {}
    for q in range(len(da)):
        d = [float(((da[q] - db[t]) ** 2).sum()) for t in range(len(db))]
        o = sorted(range(len(db)), key=lambda t: (d[t], t))
        b1, b2 = d[o[0]], d[o[1]]
        if np.sqrt(b1) < ratio * np.sqrt(b2):
            t = o[0]
            if t not in best or (b1, q) < best[t]:
                best[t] = (b1, q)
    ts = sorted(best)
    return np.array([best[t][1] for t in ts]), np.array(ts)


def test_fp64_oracle_vs_loop():
    rng = np.random.default_rng(0)
    a, b = pair_u8(rng, 60, 70, 30)
    q, t, _ = mo.match_fp64(a, b, 0.8)
    ql, tl = brute_loop(a, b, 0.8)
    np.testing.assert_array_equal(q, ql)
    np.testing.assert_array_equal(t, tl)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_f32_restatement_equals_fp64_on_u8(seed):
    rng = np.random.default_rng(seed)
    a, b = pair_u8(rng, 400, 500, 200)
    q, t = mo.ratio_match_bf_f32(a.astype(np.float32), b.astype(np.float32), 0.8)
    q2, t2, _ = mo.match_fp64(a, b, 0.8)
    np.testing.assert_array_equal(q, q2)
    np.testing.assert_array_equal(t, t2)
    assert len(q) > 100


def test_tie_rules():
    # best tie -> rejected; a unique target claimed by two queries -> smaller distance wins,
    # equal distances -> lower query index
    b = np.zeros((4, 128), np.float32)
    b[0, 0] = 10
    b[1, 0] = 10
    b[2, 1] = 100
    b[3, 2] = 100
    a = np.zeros((4, 128), np.float32)
    a[0, 0] = 10            # equidistant to b0 and b1 -> reject
    a[1, 1] = 100           # exact copy of b2
    a[2, 1] = 100           # same -> tie on target 2 -> q=1 wins
    a[3, 2] = 99            # close to b3
    q, t, _ = mo.match_fp64(a, b, 0.8)
    assert list(zip(q, t)) == [(1, 2), (3, 3)]
    q2, t2 = mo.ratio_match_bf_f32(a, b, 0.8)
    assert list(zip(q2, t2)) == [(1, 2), (3, 3)]
    # fewer than 2 on a side -> empty (matching.py:88-89)
    assert len(mo.match_fp64(a[:1], b, 0.8)[0]) == 0


def test_guided_candidates_vs_ckdtree():
    rng = np.random.default_rng(5)
    xy_b = rng.uniform(0, 500, (3000, 2))
    pred = rng.uniform(-40, 540, (800, 2))
    tree = cKDTree(xy_b)
    for kc in (1, 4, 16, 64):
        qs, cands = mo.guided_candidates(pred, xy_b, (500, 500), 30.0, kc)
        inside = np.nonzero((pred[:, 0] > -30) & (pred[:, 0] < 530) & (pred[:, 1] > -30) & (pred[:, 1] < 530))[0]
        d, idx = tree.query(pred[inside], k=kc, distance_upper_bound=30.0)
        if kc == 1:
            idx = idx[:, None]
        has = (idx < len(xy_b)).any(1)
        np.testing.assert_array_equal(qs, inside[has])
        for c, row in zip(cands, idx[has]):
            np.testing.assert_array_equal(c, np.sort(row[row < len(xy_b)]))


def test_guided_candidates_vs_ckdtree_with_exact_ties():
    """Keypoints on an integer lattice, predictions on lattice and half-lattice points, so many
    candidates lie at exactly the same distance.  The oracle's kc nearest within the radius
    (ties to the lower index) and cKDTree.query(k=kc, distance_upper_bound) (matching.py:123-131)
    agree on the strict bound and on the kc distances of every query.  Their candidate sets are
    equal wherever the kc-th and the (kc+1)-th distances differ.  On an exact tie at that
    boundary cKDTree's choice follows its tree layout: a layout detail, not an index rule."""
    rng = np.random.default_rng(8)
    gx, gy = np.meshgrid(np.arange(0, 120, 3.0), np.arange(0, 90, 3.0))
    xy_b = np.c_[gx.ravel(), gy.ravel()][rng.permutation(gx.size)]
    pred = np.c_[rng.integers(0, 40, 600) * 3.0, rng.integers(0, 30, 600) * 3.0] + rng.choice([0.0, 1.5], (600, 2))
    tree = cKDTree(xy_b)
    n_tie_boundary = 0
    for kc in (1, 3, 5, 8):
        qs, cands = mo.guided_candidates(pred, xy_b, (120, 90), 4.6, kc)
        d, idx = tree.query(pred[qs], k=kc, distance_upper_bound=4.6)
        if kc == 1:
            idx, d = idx[:, None], d[:, None]
        for q, c, row in zip(qs, cands, idx):
            row = row[row < len(xy_b)]
            dq = ((xy_b - pred[q]) ** 2).sum(1)
            np.testing.assert_array_equal(np.sort(dq[c]), np.sort(dq[row]))
            within = np.sort(dq[dq < 4.6 * 4.6])
            if len(within) > kc and within[kc - 1] == within[kc]:
                n_tie_boundary += 1
            else:
                np.testing.assert_array_equal(c, np.sort(row))
    assert n_tie_boundary > 100


def test_guided_single_candidate_always_passes():
    a = np.zeros((1, 128), np.float32)
    b = np.full((1, 128), 50, np.float32)
    q, t, _ = mo.match_fp64(a, b, 0.8, [0], [np.array([0])])
    assert list(q) == [0] and list(t) == [0]
    q2, t2 = mo.guided_f32(a, b, [0], [np.array([0])], 0.8)
    assert list(q2) == [0]


def test_guided_f32_equals_fp64_on_u8():
    rng = np.random.default_rng(9)
    a, b = pair_u8(rng, 300, 320, 150)
    xy_b = rng.uniform(0, 400, (320, 2))
    pred = rng.uniform(0, 400, (300, 2))
    qs, cands = mo.guided_candidates(pred, xy_b, (400, 400), 60.0, 0)
    q, t = mo.guided_f32(a, b, qs, cands, 0.8)
    q2, t2, _ = mo.match_fp64(a, b, 0.8, qs, cands)
    np.testing.assert_array_equal(q, q2)
    np.testing.assert_array_equal(t, t2)


def test_guided_fp64_fast_equals_reference_form():
    rng = np.random.default_rng(13)
    a, b = pair_u8(rng, 500, 600, 250)
    xy_b = rng.uniform(0, 400, (600, 2))
    xy_a = rng.uniform(-30, 430, (500, 2))
    Hm = np.array([[1.0, 0.02, 3.0], [-0.01, 1.0, -4.0], [0, 0, 1.0]])
    q, t = mo.guided_fp64_fast(a, xy_a, b, xy_b, Hm, (400, 400), 35.0, 0.8)
    pred = mo.predict(xy_a, Hm)
    qs, cands = mo.guided_candidates(pred, xy_b, (400, 400), 35.0, 0)
    q2, t2, _ = mo.match_fp64(a, b, 0.8, qs, cands)
    np.testing.assert_array_equal(q, q2)
    np.testing.assert_array_equal(t, t2)


def test_predict_matches_projective_formula():
    rng = np.random.default_rng(2)
    xy = rng.uniform(0, 100, (50, 2))
    Hm = np.array([[1.1, 0.1, 5.0], [-0.2, 0.9, 3.0], [1e-4, 2e-4, 1.0]])
    p = mo.predict(xy, Hm)
    h = np.c_[xy, np.ones(50)] @ Hm.T
    np.testing.assert_allclose(p, h[:, :2] / h[:, 2:], rtol=1e-14)


def test_device_sort_rank_rule_handles_equal_keys():
    """The merge rank rule of the device key sorts, modelled on the CPU: with many equal keys
    (e.g. sentinels) every slot is written exactly once and the output is sorted."""
    from oracle.sort_model import device_sort
    rng = np.random.default_rng(0)
    keys = [(int(a), int(b)) for a, b in zip(rng.integers(0, 5, 300), rng.integers(0, 3, 300))]
    keys += [(9, 9)] * 60
    rng.shuffle(keys)
    out = device_sort(keys, block=16)
    assert None not in out
    assert out == sorted(keys)


@pytest.mark.parametrize("seed", [0, 1])
def test_fp64_fast_equals_fp64_loop(seed):
    """The BLAS-screened FP64 oracle returns match_fp64's list exactly, on SIFT-like float
    descriptors with exact duplicates (equal distances), exact copies of queries (distance 0)
    and on uint8 RootSIFT-like rows (many equal integer distances)."""
    from fieldstitch_amd import synth
    rng = np.random.default_rng(seed)
    da, db, _ = synth.descriptor_pair(rng, 700, 900)
    db[100] = db[200]
    db[300] = da[5]
    db[301] = da[6]
    db[302] = da[6]
    da[7] = da[8]
    for a, b in ((da, db), (np.rint(da / 4), np.rint(db / 4)), (da[:3], db), (da, db[:2])):
        slow = mo.match_fp64(a, b, 0.8)
        fast = mo.match_fp64_fast(a, b, 0.8, block=256)
        for x, y in zip(slow, fast):
            np.testing.assert_array_equal(x, y)


def test_guided_f32_fast_equals_fp64_on_u8():
    """The vectorised float32 guided matcher (the CPU baseline) returns the FP64 oracle's list
    on uint8 RootSIFT-like descriptors with planted matches (exact integer distances)."""
    rng = np.random.default_rng(12)
    nb, na = 3000, 2500
    xy_b = rng.uniform(0, 800, (nb, 2))
    d = rng.gamma(0.6, 1.0, (nb, 128))
    db = np.clip(np.sqrt(d / d.sum(1, keepdims=True)) * 512, 0, 255).astype(np.uint8)
    src = rng.integers(0, nb, na)
    xy_a = xy_b[src] - (20.0, 10.0) + rng.normal(0, 2.0, (na, 2))
    da = np.clip(db[src].astype(int) + rng.integers(-9, 10, (na, 128)), 0, 255).astype(np.uint8)
    Hm = np.array([[1.0, 0, 20.0], [0, 1.0, 10.0], [0, 0, 1.0]])
    q, t = mo.guided_f32_fast(da, xy_a, db, xy_b, Hm, (800, 800), 25.0, 0.8, chunk=300)
    rq, rt = mo.guided_fp64_fast(da, xy_a, db, xy_b, Hm, (800, 800), 25.0, 0.8)
    assert len(rq) > 500
    np.testing.assert_array_equal(q, rq)
    np.testing.assert_array_equal(t, rt)
