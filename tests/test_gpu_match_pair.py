"""match_pair (matching.py:205-253) end to end on a synthetic 3-photo strip: feature files in
the reference's npz format, the worker context, the guided path, the exhaustive fallback and
the failure tuples; the verified matches are checked against the true homographies."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def strip(tmp_path_factory, gpu):
    import torch
    from fieldstitch_amd import features, synth
    d = tmp_path_factory.mktemp("feat")
    W, H = 1320, 989
    fl = synth.flight(seed=21, rows=1, cols=3, W=W, H=H, workers=8)
    fb = features.FeatureBatch([im.shape for im in fl.images], nfeat=3000, contrast=0.04, edge=10.0)
    fb.run([torch.from_numpy(im).cuda() for im in fl.images])
    fb.check_capacity()
    n = len(fl.images)
    for i in range(n):
        xy, u8, resp, _ = fb.results(i)
        features._save(str(d / ("%d.npz" % i)), xy, None, resp, W, H, desc_u8=u8)
    features._save(str(d / ("%d.npz" % n)), xy[:12], None, resp[:12], W, H, desc_u8=u8[:12])   # too few
    m = n + 1
    K = np.array([[1000.0, 0.0, W / 2.0], [0.0, 1000.0, H / 2.0], [0.0, 0.0, 1.0]])
    ctx = {"feat_dir": str(d), "W": [W] * m, "H": [H] * m, "K": [K] * m, "dist": [np.zeros(5)] * m,
           "R": [np.eye(3)] * m, "C": [np.zeros(3)] * m, "zg": [0.0] * m, "has_rot": [False] * m,
           "prior_tol": 0.18, "ransac_thr": 1.0, "min_inliers": 20, "ratio": 0.8, "checks": 48, "use_prior": True,
           "guided": True, "n_coarse": 1000, "guided_radius": 0.03, "guided_k": 16}
    return fl, ctx


def _check(fl, res, i, j, min_in=40):
    from fieldstitch_amd import matching
    ri, rj, q, t, sim, n_in = res
    assert (ri, rj) == (i, j)
    assert q is not None and q.dtype == np.int32 and t.dtype == np.int32 and len(q) == len(t) == n_in >= min_in
    assert len(np.unique(t)) == len(t)
    A, B = matching._get(i), matching._get(j)
    Ht = fl.H_true[(i, j)]
    p = np.c_[A["und"][q], np.ones(len(q))] @ Ht.T
    err = np.linalg.norm(p[:, :2] / p[:, 2:] - B["und"][t], axis=1)
    assert (err < 3.0).mean() > 0.98, np.percentile(err, [50, 99])
    ang, sc, tx, ty = sim
    assert abs(ang - np.arctan2(Ht[1, 0], Ht[0, 0])) < 0.01
    assert abs(sc / np.hypot(Ht[0, 0], Ht[1, 0]) - 1.0) < 0.01
    return n_in


def test_match_pair_guided(strip):
    from fieldstitch_amd import matching
    fl, ctx = strip
    matching.init_worker(ctx)
    for i, j in ((0, 1), (1, 2)):
        n_in = _check(fl, matching.match_pair((i, j)), i, j)
        print("pair", i, j, "verified", n_in, flush=True)


def test_match_pair_fallback_and_failures(strip):
    from fieldstitch_amd import matching
    fl, ctx = strip
    matching.init_worker(dict(ctx, guided=False))
    _check(fl, matching.match_pair((0, 1)), 0, 1, min_in=20)
    assert matching.match_pair((0, 3)) == (0, 3, None, None, None, 0)            # < 30 features
    i, j, q, t, sim, n = matching.match_pair((0, 99))                            # no feature file
    assert (i, j, q, t, n) == (0, 99, None, None, 0) and sim[0] == "error"
