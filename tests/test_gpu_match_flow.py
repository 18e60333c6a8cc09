"""The batched match_pair chain (match_pairs) on a 2 x 3 synthetic flight with nadir cameras
(ground-plane prior on): every pair's tuple equals the CPU restatement of matching.py:205-253
on the oracle's exact matchers and RANSAC (q, t and the inlier count exactly, the similarity to
1e-9), whatever batch the pair runs in; the guided path, the fallback and the failure tuples."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CTX = {"prior_tol": 0.18, "ransac_thr": 1.0, "min_inliers": 20, "ratio": 0.8, "checks": 48, "use_prior": True,
       "guided": True, "n_coarse": 1000, "guided_radius": 0.03, "guided_k": 16}


@pytest.fixture(scope="module")
def flight(tmp_path_factory, gpu):
    import torch
    from fieldstitch_amd import features, synth
    d = tmp_path_factory.mktemp("flow")
    fl = synth.flight(seed=23, rows=2, cols=3, W=1320, H=989, workers=1)
    shapes = [im.shape for im in fl.images]
    fb = features.FeatureBatch(shapes, nfeat=3000, contrast=0.04, edge=10.0)
    fb.run([torch.from_numpy(im).cuda() for im in fl.images])
    fb.check_capacity()
    feats = {}
    for i, (h, w) in enumerate(shapes):
        xy, u8, resp, _ = fb.results(i)
        size, angle = fb.size_angle(i)
        features._save(str(d / ("%d.npz" % i)), xy, None, resp, w, h, desc_u8=u8, size=size, angle=angle)
        feats[i] = (u8, xy.astype(np.float64))     # no distortion, reduce 1: und = xy
    ctx = dict(CTX, feat_dir=str(d), **synth.flight_cameras(fl, seed=3))
    return fl, ctx, feats


def _same(got, want):
    assert got[:2] == want[:2]
    assert got[5] == want[5], (got[:2], got[5], want[5])
    if want[2] is None:
        assert got[2] is None and got[3] is None and got[4] is None
        return
    assert got[2].dtype == np.int32 and got[3].dtype == np.int32
    np.testing.assert_array_equal(got[2], want[2])
    np.testing.assert_array_equal(got[3], want[3])
    if want[4] is None:
        assert got[4] is None
    else:
        np.testing.assert_allclose(got[4], want[4], rtol=1e-9, atol=1e-7)


@pytest.mark.parametrize("variant", ["guided", "no_prior", "fallback"])
def test_match_pairs_equals_chain_oracle(flight, variant):
    from fieldstitch_amd import matching
    from oracle import match_pair_oracle as mpo
    fl, ctx, feats = flight
    ctx = dict(ctx)
    if variant == "no_prior":
        ctx["use_prior"] = False
    if variant == "fallback":
        ctx["guided"] = False
    matching.init_worker(ctx)
    pairs = list(fl.pairs)
    stats = {}
    got = matching.match_pairs(pairs, stats=stats)
    print("variant", variant, stats, flush=True)
    assert stats["matched"] >= len(pairs) - 2
    if variant != "fallback":
        assert stats["guided_verified"] >= len(pairs) // 2
    for (i, j), g in zip(pairs, got):
        _same(g, mpo.match_pair(ctx, feats, i, j))


def test_batch_independence_and_failures(flight):
    """A pair's tuple is the same alone, in a batch, and in another order; a pair with a missing
    feature file reports an error tuple without disturbing the others."""
    from fieldstitch_amd import matching
    fl, ctx, _ = flight
    matching.init_worker(ctx)
    pairs = list(fl.pairs)
    batch = matching.match_pairs(pairs)
    rev = matching.match_pairs(pairs[::-1])[::-1]
    for p, a, b in zip(pairs, batch, rev):
        _same(a, b)
        _same(matching.match_pair(p), a)
    mixed = matching.match_pairs([pairs[0], (0, 99), pairs[1]])
    _same(mixed[0], batch[0])
    _same(mixed[2], batch[1])
    i, j, q, t, sim, n = mixed[1]
    assert (i, j, q, t, n) == (0, 99, None, None, 0) and sim[0] == "error"
