"""Batched execution: graph replays of the feature pipeline, launch-shape tunables (results
must not depend on them) and the autotuner."""
import numpy as np
import pytest

from fieldstitch_amd import synth

pytestmark = pytest.mark.gpu


def _images(seed, shapes):
    rng = np.random.default_rng(seed)
    out = []
    for k, (H, W) in enumerate(shapes):
        img1, _, _ = synth.field_pair(seed + k, W=max(W, 64), H=max(H, 64), n_poly=12)
        out.append(np.ascontiguousarray(img1[:H, :W]))
    return out


def _snapshot(fb):
    n_det, n_sel = fb.counts()
    res = [fb.results(i) for i in range(fb.n)]
    return n_det, [(r[0].copy(), r[1].copy(), r[2].copy()) for r in res]


def _same(a, b):
    if not np.array_equal(a[0], b[0]):
        return False
    return all(np.array_equal(x, y) for ra, rb in zip(a[1], b[1]) for x, y in zip(ra, rb))


SHAPES = [(240, 320), (300, 416), (240, 320), (300, 416)]


def test_feature_graph_replays_equal_eager(gpu):
    import torch
    from fieldstitch_amd.features import FeatureBatch
    hosts = _images(3, SHAPES)
    dev = [torch.from_numpy(h).cuda() for h in hosts]
    fb = FeatureBatch(SHAPES, nfeat=600, contrast=0.04, edge=10.0)
    fb.run(dev)
    fb.check_capacity()
    ref = _snapshot(fb)
    for rep in range(3):
        fb.run(dev, graph=True)
        torch.cuda.synchronize()
        assert _same(_snapshot(fb), ref), "replay %d differs" % rep
        fb.run(dev)                       # eager work in between
        torch.cuda.synchronize()
    # new photos into the same buffers, then replay: equals an eager run on them
    new = _images(11, SHAPES)
    for d, h in zip(dev, new):
        d.copy_(torch.from_numpy(h))
    fb.run(dev, graph=True)
    torch.cuda.synchronize()
    got = _snapshot(fb)
    fb.run(dev)
    torch.cuda.synchronize()
    assert _same(got, _snapshot(fb))
    assert not _same(got, ref)


def test_tunables_do_not_change_results(gpu):
    import torch
    from fieldstitch_amd import autotune
    from fieldstitch_amd.features import FeatureBatch
    hosts = _images(5, SHAPES)
    dev = [torch.from_numpy(h).cuda() for h in hosts]
    fb = FeatureBatch(SHAPES, nfeat=600, contrast=0.04, edge=10.0)
    keys = (autotune.ORIENT_BLOCKS, autotune.REFINE_BLOCKS, autotune.BLUR_ROWS)
    old = [autotune.get(k) for k in keys]
    snaps = []
    try:
        for vals in ((1024, 64, 8), (8192, 512, 32), (2048, 128, 24), (4096, 256, 16)):
            for k, v in zip(keys, vals):
                autotune.set(k, v)
            print("tunables", vals, flush=True)
            fb.run(dev)
            torch.cuda.synchronize()
            snaps.append(_snapshot(fb))
    finally:
        for k, v in zip(keys, old):
            autotune.set(k, v)
    assert all(_same(snaps[0], sn) for sn in snaps[1:])


def test_autotune_features_and_float_matcher(gpu):
    import torch
    from fieldstitch_amd import autotune
    from fieldstitch_amd.features import FeatureBatch
    from fieldstitch_amd.matching import FloatMatcher
    hosts = _images(7, SHAPES)
    dev = [torch.from_numpy(h).cuda() for h in hosts]
    fb = FeatureBatch(SHAPES, nfeat=600, contrast=0.04, edge=10.0)
    out = autotune.tune_features(fb, dev, reps=2)
    assert out["chosen"]["orientation_blocks"] in autotune.CANDIDATES[autotune.ORIENT_BLOCKS]
    assert out["chosen"]["refine_blocks"] in autotune.CANDIDATES[autotune.REFINE_BLOCKS]
    assert out["chosen"]["blur_rows"] in autotune.CANDIDATES[autotune.BLUR_ROWS]
    rng = np.random.default_rng(2)
    da, db, _ = synth.descriptor_pair(rng, 900, 1100)
    fm = FloatMatcher([da, db], [(0, 1)], ratio=0.8, precision="bf16")
    res = {}
    for qb in (1, 2):
        for tile in (64, 128):
            autotune.set(autotune.FSCREEN_QB, qb)
            autotune.set(autotune.FSCREEN_TILE, tile)
            fm.run()
            res[(qb, tile)] = fm.results()[0]
    ref = res[(1, 64)]
    assert all(np.array_equal(r[0], ref[0]) and np.array_equal(r[1], ref[1]) for r in res.values())
    t = autotune.tune_float_matcher(fm, reps=2)
    assert t["chosen"] in (64, 128) and t["chosen_query_blocks"] in (1, 2)
