"""The CLAHE restatement (cv2.createCLAHE(2.0, (8, 8)).apply, features.py:70-71) against what
the mathematics fixes; parity with cv2 itself is unpinned (cv2 is not installed)."""
import numpy as np

from oracle import clahe_oracle as O


def test_one_tile_without_clip_is_histogram_equalisation():
    rng = np.random.default_rng(0)
    img = rng.integers(0, 256, (64, 48)).astype(np.uint8)
    out = O.apply(img, clip=0.0, tiles_x=1, tiles_y=1)
    cdf = np.cumsum(np.bincount(img.ravel(), minlength=256))
    expect = np.rint(cdf.astype(np.float64) * 255.0 / img.size)
    assert np.abs(out.astype(np.int64) - expect[img]).max() == 0


def test_clip_and_redistribution_by_hand():
    h = np.zeros(256, np.int64)
    h[10] = 100
    h[20] = 7
    # limit 5: excess 95 + 2 = 97 -> every bin +0, residual 97 -> bins 0, 2, 4, ... (step 2) get +1
    c = O.clip_histogram(h, 5)
    assert c.sum() == h.sum()
    assert c[10] == 6 and c[20] == 6 and c[1] == 0 and c[2] == 1 and c[192] == 1 and c[194] == 0
    h = np.full(256, 3, np.int64)
    h[0] = 3 + 600
    c = O.clip_histogram(h, 4)         # excess 599: +2 everywhere, residual 87, step 2
    assert c.sum() == h.sum() and c[0] == 7 and c[1] == 5 and c[172] == 6 and c[174] == 5


def test_lut_invariants_and_padding():
    rng = np.random.default_rng(1)
    img = (rng.normal(90, 30, (101, 133)).clip(0, 255)).astype(np.uint8)
    L, (th, tw) = O.luts(img)
    assert (th, tw) == (13, 17)                       # 101 + 3 = 104, 133 + 3 = 136
    assert np.all(L[..., 255] == 255) and np.all(np.diff(L.astype(int), axis=-1) >= 0)
    ext = O.extend(img)
    assert ext[101, 5] == img[99, 5] and ext[5, 134] == img[5, 130]     # BORDER_REFLECT_101
    img2 = img[:96, :128]                              # divisible: no extension
    assert O.extend(img2) is img2
    img3 = img[:96, :130]                              # rows divide, columns do not: 8 extra rows
    assert O.extend(img3).shape == (104, 136)


def test_identical_tiles_map_through_the_tile_lut():
    rng = np.random.default_rng(2)
    tile = rng.integers(0, 256, (12, 10)).astype(np.uint8)
    img = np.tile(tile, (8, 8))
    L, _ = O.luts(img)
    assert np.all(L == L[0, 0])
    assert np.array_equal(O.apply(img), L[0, 0][img])


def test_constant_image():
    img = np.full((80, 64), 77, np.uint8)
    out = O.apply(img)
    assert np.all(out == out[0, 0])
