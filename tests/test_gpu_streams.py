"""The drop-in calls on a caller's side stream: gsift.detect / describe, CLAHE.apply and
build_tracks give the same results on a non-default stream as on the default one (every
buffer, launch and copy of a call is ordered on the stream it was given)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _photo(seed, H=480, W=640):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:H, 0:W]
    img = 120 + 60 * np.sin(x / 9.0) * np.cos(y / 13.0) + rng.normal(0, 12, (H, W))
    return img.clip(0, 255).astype(np.uint8)


def test_sift_on_side_stream(gpu):
    import torch
    from fieldstitch_amd import gsift
    img = _photo(1)
    k0, s0 = gsift.detect(img, max_kp=4000, contrast=0.04, edge=10.0)
    d0 = gsift.describe(s0, k0, np.arange(len(k0["x"])))
    side = torch.cuda.Stream()
    print("sift side stream", len(k0["x"]), flush=True)
    k1, s1 = gsift.detect(torch.from_numpy(img).cuda(), max_kp=4000, contrast=0.04, edge=10.0, stream=side)
    d1, u1 = gsift.describe(s1, k1, np.arange(len(k1["x"])), rootsift_u8=True)
    assert len(k0["x"]) > 100
    for key in k0:
        assert np.array_equal(k0[key], k1[key]), key
    assert np.array_equal(d0, d1) and u1.shape == d1.shape


def test_clahe_on_side_stream(gpu):
    import torch
    from fieldstitch_amd import clahe
    imgs = np.stack([_photo(s, 200, 256) for s in (2, 3)])
    c = clahe.createCLAHE()
    want = c.apply(torch.from_numpy(imgs).cuda()).cpu().numpy()
    side = torch.cuda.Stream()
    src = torch.from_numpy(imgs).cuda()
    out = torch.empty_like(src)
    got = c.apply(src, out=out, stream=side)
    side.synchronize()
    assert got.data_ptr() == out.data_ptr()
    assert np.array_equal(out.cpu().numpy(), want)
    assert np.array_equal(c.apply(imgs[0], stream=side), want[0])
    with pytest.raises(ValueError):
        c.apply(src, out=torch.empty((2, 200, 255), dtype=torch.uint8, device="cuda"))
    with pytest.raises(ValueError):
        c.apply(src, out=torch.empty((2, 200, 256), dtype=torch.int16, device="cuda"))


def test_tracks_on_side_stream(gpu):
    import torch
    from fieldstitch_amd import tracks
    rng = np.random.default_rng(5)
    nfeat = [300, 250, 280, 310]
    pairs = []
    for i in range(4):
        for j in range(i + 1, 4):
            m = 120
            pairs.append((i, j, rng.integers(0, nfeat[i], m), rng.integers(0, nfeat[j], m)))
    want = tracks.build_tracks(pairs, nfeat, 4)
    got = tracks.build_tracks(pairs, nfeat, 4, stream=torch.cuda.Stream())
    assert want[3] == got[3] and all(np.array_equal(a, b) for a, b in zip(want[:3], got[:3]))
