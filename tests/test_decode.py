"""Photo decoding for the feature worker (features.py:18-26 decode, 60-71 prepare): JPEG photos
reduced by 1, 2, 4 or 8 in the DCT domain (libjpeg through Pillow), luma only; the thumbnail
of prepare.  cv2 is not installed, so equality with cv2.imdecode is unpinned; the checks are
the output sizes libjpeg's scaled decode defines, the exact gray plane of a gray JPEG, the
block averages a DCT-domain reduction approximates (exact DC average at 1/8), and the error
cases."""
import numpy as np
import pytest

PIL = pytest.importorskip("PIL")


def _photo(h, w, seed=0):
    rng = np.random.default_rng(seed)
    y, x = np.mgrid[0:h, 0:w]
    base = 120 + 60 * np.sin(x / 23.0) * np.cos(y / 31.0)
    rgb = np.stack([base + 30 * np.sin(y / 7.0), base, base - 25 * np.cos(x / 11.0)], -1)
    for _ in range(12):
        cy, cx, r = rng.integers(0, h), rng.integers(0, w), rng.integers(5, 40)
        rgb[(y - cy) ** 2 + (x - cx) ** 2 < r * r] += rng.uniform(-60, 60, 3)
    return np.clip(rgb + rng.normal(0, 3, rgb.shape), 0, 255).astype(np.uint8)


def _save(path, arr, fmt="JPEG", **kw):
    from PIL import Image
    Image.fromarray(arr).save(path, format=fmt, **kw)
    return str(path)


def _block_mean(g, r):
    h, w = g.shape
    hp, wp = -(-h // r) * r, -(-w // r) * r
    pad = np.pad(g.astype(np.float64), ((0, hp - h), (0, wp - w)), mode="edge")
    return pad.reshape(hp // r, r, wp // r, r).mean((1, 3))


def test_gray_jpeg_full_and_reduced(tmp_path):
    from PIL import Image
    from fieldstitch_amd.features import read_gray
    h, w = 301, 457
    rgb = _photo(h, w)
    gray = np.asarray(Image.fromarray(rgb).convert("L"))
    p = _save(tmp_path / "g.jpg", gray, quality=95)
    full = read_gray(p, 1)
    assert full.dtype == np.uint8 and full.shape == (h, w)
    np.testing.assert_array_equal(full, np.asarray(Image.open(p)))
    for r in (2, 4, 8):
        g = read_gray(p, r)
        assert g.shape == (-(-h // r), -(-w // r)) and g.flags["C_CONTIGUOUS"]
        d = np.abs(g.astype(np.float64) - _block_mean(full, r))
        # interior blocks (the edge blocks of a partial 8x8 MCU are padded by the codec)
        di = d[: h // r - 1, : w // r - 1]
        assert di.mean() < 0.6 and di.max() < 2.0, (r, di.mean(), di.max())


def test_colour_jpeg_gray_is_its_luma(tmp_path):
    from PIL import Image
    from fieldstitch_amd.features import read_gray
    h, w = 240, 320
    p = _save(tmp_path / "c.jpg", _photo(h, w, 3), quality=95)
    g = read_gray(p, 1)
    luma = np.asarray(Image.open(p).convert("RGB").convert("L")).astype(int)
    assert g.shape == (h, w)
    assert np.abs(g.astype(int) - luma).max() <= 3


def test_decode_errors_and_other_formats(tmp_path):
    from fieldstitch_amd.features import read_gray
    rgb = _photo(64, 80, 5)
    png = _save(tmp_path / "p.png", rgb, fmt="PNG")
    assert read_gray(png, 1).shape == (64, 80)
    with pytest.raises(ValueError, match="reduced decoding is for JPEG"):
        read_gray(png, 2)
    jpg = _save(tmp_path / "j.jpg", rgb)
    with pytest.raises(ValueError, match="reduce must be"):
        read_gray(jpg, 3)
    with pytest.raises(FileNotFoundError):
        read_gray(str(tmp_path / "missing.jpg"), 1)
    bad = tmp_path / "bad.jpg"
    bad.write_bytes(b"not an image at all")
    with pytest.raises(Exception, match="cannot identify"):
        read_gray(str(bad), 1)
    npy = tmp_path / "a.npy"
    np.save(npy, rgb[..., 0])
    np.testing.assert_array_equal(read_gray(str(npy), 4), rgb[..., 0])    # stored already reduced


def test_prepare_task_writes_the_thumbnail(tmp_path):
    from PIL import Image
    from fieldstitch_amd.features import prepare_task, read_gray
    h, w = 301, 457
    p = _save(tmp_path / "c.jpg", _photo(h, w, 7), quality=92)
    for r in (1, 2, 8):
        thumb = str(tmp_path / ("t%d.jpg" % r))
        g = prepare_task((0, p, str(tmp_path / "f.npz"), thumb, r, 1000))
        assert g.shape == (-(-h // r), -(-w // r)) and g.dtype == np.uint8
        with Image.open(thumb) as t:
            col_w, col_h = -(-w // r), -(-h // r)
            exp = (col_w, col_h) if r == 8 else ((col_w + 1) * r // 8, (col_h + 1) * r // 8)
            assert t.size == exp and t.format == "JPEG"
        # the gray image is the colour decode's luma: close to the direct gray decode
        assert np.abs(g.astype(int) - read_gray(p, r).astype(int)).mean() < 1.5
        # thumbnail present: the gray image is decoded directly
        np.testing.assert_array_equal(prepare_task((0, p, None, thumb, r, 1000)), read_gray(p, r))
