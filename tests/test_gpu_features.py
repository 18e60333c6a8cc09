"""GPU feature batch (detect -> device grid spread -> describe -> RootSIFT u8 in one call)
against the drop-in per-image path (host _spread of features.py) and the CPU oracle."""
import numpy as np
import pytest

from fieldstitch_amd import synth

pytestmark = pytest.mark.gpu


def per_image_path(img, nfeat, contrast=0.015, edge=12.0):
    """The drop-in calls composed photo by photo: gsift.detect with 2.2 nfeat keypoints, the
    host grid spread, a stable response sort, gsift.describe -> (xy, desc, resp)."""
    from fieldstitch_amd import features, gsift
    k, st = gsift.detect(img, max_kp=int(nfeat * 2.2), contrast=contrast, edge=edge)
    if len(k["x"]) < 10:
        return np.zeros((0, 2), np.float32), None, np.zeros(0, np.float32)
    sel = features.spread_select(k["x"], k["y"], k["response"], img.shape[1], img.shape[0], nfeat)
    sel = sel[np.argsort(-k["response"][sel], kind="stable")]
    d = gsift.describe(st, k, sel)
    return np.stack([k["x"][sel], k["y"][sel]], 1).astype(np.float32), d, k["response"][sel].astype(np.float32)


def _images():
    img1, img2, _ = synth.field_pair(1)
    rng = np.random.default_rng(21)
    mos = synth.Mosaic(seed=int(rng.integers(1 << 30)))
    mos.polys = synth.random_polygons(rng, 10, 0, 0, 500, 400)
    img3 = mos.render(np.eye(3)[:2], 400, 500, rng)
    return [img1, img2, img3, img1[:600, :900].copy()]


@pytest.mark.parametrize("nfeat", [1000, 2500, 20000])
def test_feature_batch_equals_per_image_path(gpu, nfeat):
    import torch
    from fieldstitch_amd import features
    imgs = _images()
    print("case nfeat", nfeat, [i.shape for i in imgs], flush=True)
    fb = features.FeatureBatch([i.shape for i in imgs], nfeat=nfeat, contrast=0.015, edge=12.0, float_desc=True)
    dev = [torch.from_numpy(i).cuda() for i in imgs]
    fb.run(dev)
    fb.check_capacity()
    for i, img in enumerate(imgs):
        xy, u8, resp, d = fb.results(i)
        # drop-in path: detect with max_kp = 2.2 nfeat, host spread, stable response sort
        xr, dr, rr = per_image_path(img, nfeat)
        assert xy.shape == xr.shape, (i, xy.shape, xr.shape)
        np.testing.assert_array_equal(xy, xr)
        np.testing.assert_array_equal(resp, rr)
        np.testing.assert_array_equal(d, dr)
        np.testing.assert_array_equal(u8, features.rootsift_u8(dr))


def test_spread_matches_reference_formula(gpu):
    """Device spread == host _spread on a case with many keypoints per cell and equal
    responses (multi-orientation keypoints)."""
    import torch
    from fieldstitch_amd import features, gsift
    img = _images()[0]
    nfeat = 700
    kps, st = gsift.detect(img, max_kp=int(nfeat * 2.2), contrast=0.015, edge=12.0)
    sel = features.spread_select(kps["x"], kps["y"], kps["response"], img.shape[1], img.shape[0], nfeat)
    sel = sel[np.argsort(-kps["response"][sel], kind="stable")]
    fb = features.FeatureBatch([img.shape], nfeat=nfeat, contrast=0.015, edge=12.0)
    fb.run([torch.from_numpy(img).cuda()])
    np.testing.assert_array_equal(fb._sel(0, nfeat), sel)
    assert len(np.unique(kps["response"])) < len(kps["response"])   # ties present


def test_sift_gpu_drop_in_equals_per_image_path(gpu):
    """features.sift_gpu (a one-photo FeatureBatch) returns the per-photo composition's arrays
    and the reference's empty result for a photo with fewer than 10 keypoints."""
    from fieldstitch_amd import features
    img = _images()[2]
    xy, d, resp, t = features.sift_gpu(img, 900)
    xr, dr, rr = per_image_path(img, 900)
    np.testing.assert_array_equal(xy, xr)
    np.testing.assert_array_equal(d, dr)
    np.testing.assert_array_equal(resp, rr)
    assert t >= 0.0
    flat = np.full((64, 80), 128, np.uint8)
    xy, d, resp, _ = features.sift_gpu(flat, 900)
    assert xy.shape == (0, 2) and d is None and resp.shape == (0,)


def test_feature_batch_full_size(gpu):
    """Two photos of the benchmark's size (2640x1978, nfeat 8000) through one FeatureBatch
    equal the per-image drop-in path (detect with its default contrast 0.015 and edge 12,
    host _spread, sort, describe, RootSIFT) bit for bit."""
    import torch
    from fieldstitch_amd import features
    fl = synth.flight(seed=2, rows=1, cols=2, workers=8)
    imgs = fl.images
    print("case full size", [i.shape for i in imgs], flush=True)
    fb = features.FeatureBatch([i.shape for i in imgs], nfeat=8000, contrast=0.015, edge=12.0, float_desc=True)
    fb.run([torch.from_numpy(i).cuda() for i in imgs])
    fb.check_capacity()
    for i, img in enumerate(imgs):
        xy, u8, resp, d = fb.results(i)
        xr, dr, rr = per_image_path(img, 8000)
        assert len(xy) == 8000
        np.testing.assert_array_equal(xy, xr)
        np.testing.assert_array_equal(resp, rr)
        np.testing.assert_array_equal(d, dr)
        np.testing.assert_array_equal(u8, features.rootsift_u8(dr))
