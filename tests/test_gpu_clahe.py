"""GPU CLAHE (fieldstitch_amd.clahe, features.py:60-71) against the CPU restatement of
OpenCV's algorithm: bit-exact on divisible, non-divisible and half-divisible sizes, several
clip limits and grids, batches and in place."""
import numpy as np
import pytest

from oracle import clahe_oracle as O

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("H,W", [(96, 128), (101, 133), (96, 130), (1978, 2640), (37, 9)])
@pytest.mark.parametrize("clip,grid", [(2.0, (8, 8)), (0.0, (8, 8)), (40.0, (4, 6)), (2.0, (1, 1)), (0.5, (3, 2))])
def test_clahe_equals_oracle(gpu, H, W, clip, grid):
    from fieldstitch_amd import clahe
    if W <= grid[0] or H <= grid[1]:
        pytest.skip("image smaller than the grid")
    rng = np.random.default_rng(H * 7 + W)
    img = (rng.normal(100, 40, (H, W)) + 30 * np.sin(np.arange(W) / 17.0)[None, :]).clip(0, 255).astype(np.uint8)
    print("clahe H=%d W=%d clip=%g grid=%s" % (H, W, clip, grid), flush=True)
    out = clahe.createCLAHE(clip, grid).apply(img)
    assert np.array_equal(out, O.apply(img, clip, grid[0], grid[1]))


def test_batch_and_in_place(gpu):
    import torch
    from fieldstitch_amd import clahe
    rng = np.random.default_rng(3)
    imgs = rng.integers(0, 256, (3, 120, 161)).astype(np.uint8)
    d = torch.from_numpy(imgs).cuda()
    c = clahe.createCLAHE()
    out = c.apply(d).cpu().numpy()
    for k in range(3):
        assert np.array_equal(out[k], O.apply(imgs[k]))
    c.apply(d, out=d)                        # in place
    assert np.array_equal(d.cpu().numpy(), out)
    # width divisible by 4 and tiles >= 16 rows: the LDS-staged 4-pixel form, in place too
    imgs = rng.integers(0, 256, (2, 200, 256)).astype(np.uint8)
    d = torch.from_numpy(imgs).cuda()
    c.apply(d, out=d)
    for k in range(2):
        assert np.array_equal(d[k].cpu().numpy(), O.apply(imgs[k]))
    with pytest.raises(TypeError):
        c.apply(imgs[0].astype(np.float32))
