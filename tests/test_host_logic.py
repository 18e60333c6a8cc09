"""Host-side logic of the drop-ins, checked on the CPU against independent formulations: the
grid spread (features.py:38-57) against a per-cell loop over the definition, the
one-query-per-target rule (matching.py:103-110) against the oracle's sort form, the RootSIFT
encoding (features.py:80-82) against the C oracle, the similarity fit against a least-squares
solve, and the RANSAC seeds."""
import numpy as np
import pytest


def spread_by_loops(x, y, resp, W, H, n, gx=12, gy=9):
    """The spread, written cell by cell: strongest first inside a cell (ties by index), the
    cells in row-major order, each giving at most ceil(n / (gx gy)); then the strongest of the
    rest, then the first n."""
    m = len(resp)
    if m <= n:
        return list(range(m))
    quota = int(np.ceil(n / (gx * gy)))
    cells = {}
    for k in range(m):
        cx = min(max(int(np.float32(x[k]) / np.float32(W) * np.float32(gx)), 0), gx - 1)
        cy = min(max(int(np.float32(y[k]) / np.float32(H) * np.float32(gy)), 0), gy - 1)
        cells.setdefault(cy * gx + cx, []).append(k)
    out = []
    for c in sorted(cells):
        members = sorted(cells[c], key=lambda k: (-float(resp[k]), k))
        out += members[:quota]
    if len(out) < n:
        used = set(out)
        rest = sorted((k for k in range(m) if k not in used), key=lambda k: (-float(resp[k]), k))
        out += rest[:n - len(out)]
    return out[:n]


@pytest.mark.parametrize("seed,m,n", [(0, 3000, 700), (1, 500, 450), (2, 2000, 108), (3, 50, 60), (4, 4000, 3900)])
def test_spread_select_equals_cell_loops(seed, m, n):
    from fieldstitch_amd import features
    rng = np.random.default_rng(seed)
    W, H = 1320, 989
    # clustered points (empty and crowded cells) with many equal responses
    x = np.concatenate([rng.uniform(0, W, m // 2), rng.normal(300, 40, m - m // 2)]).astype(np.float32)
    y = np.concatenate([rng.uniform(0, H, m // 2), rng.normal(200, 30, m - m // 2)]).astype(np.float32)
    x[:5] = [-3.0, W, W + 10, 0.0, W - 1e-3]
    resp = np.round(rng.gamma(1.0, 0.01, m), 3).astype(np.float32)
    got = features.spread_select(x, y, resp, W, H, n)
    assert got.tolist() == spread_by_loops(x, y, resp, W, H, n)
    assert len(np.unique(resp)) < m

    class K:
        def __init__(self, px, py, r):
            self.pt, self.response = (float(px), float(py)), float(r)
    assert features._spread([K(a, b, r) for a, b, r in zip(x, y, resp)], W, H, n) == got.tolist()


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_unique_equals_oracle(seed):
    from fieldstitch_amd import matching
    from oracle import match_oracle as mo
    rng = np.random.default_rng(seed)
    n = 3000
    q = np.arange(n, dtype=np.int64)
    t = rng.integers(0, 700, n).astype(np.int64)
    dd = rng.integers(0, 40, n).astype(np.float32)      # many equal distances per target
    a = matching._unique(q, t, dd)
    b = mo.unique(q, t, dd)
    np.testing.assert_array_equal(a[0], b[0])
    np.testing.assert_array_equal(a[1], b[1])
    assert np.all(np.diff(a[1]) > 0)
    e = matching._unique(q[:0], t[:0], dd[:0])
    assert len(e[0]) == 0 and len(e[1]) == 0


def test_rootsift_equals_c_oracle():
    from fieldstitch_amd import features, synth
    from oracle import sift_oracle as so
    rng = np.random.default_rng(4)
    d = synth.sift_like_descriptors(rng, 2000)
    d[0] = 0.0
    np.testing.assert_array_equal(features.rootsift_u8(d), so.rootsift_u8(d))


def test_save_format(tmp_path):
    from fieldstitch_amd import features
    rng = np.random.default_rng(5)
    xy = rng.uniform(0, 100, (40, 2)).astype(np.float32)
    d = rng.uniform(0, 50, (40, 128)).astype(np.float32)
    r = rng.uniform(0, 1, 40).astype(np.float32)
    p = str(tmp_path / "a.npz")
    assert features._save(p, xy, d, r, 320, 240, size=np.ones(40), angle=np.zeros(40)) == 40
    f = np.load(p)
    assert f["xy"].dtype == np.float32 and f["desc"].dtype == np.uint8 and f["red_wh"].tolist() == [320, 240]
    np.testing.assert_array_equal(f["desc"], features.rootsift_u8(d))
    assert f["size"].dtype == np.float32 and len(f["angle"]) == 40
    assert features._save(p, xy[:9], d[:9], r[:9], 320, 240) == 0
    f = np.load(p)
    assert f["xy"].shape == (0, 2) and f["desc"].shape == (0, 128) and "size" not in f.files
    assert features._save(p, xy, None, r, 320, 240) == 0


def test_similarity_fit_equals_least_squares():
    from fieldstitch_amd import match_flow
    rng = np.random.default_rng(6)
    a = rng.uniform(0, 2000, (500, 2))
    th, s, t = 0.3, 1.07, np.array([120.0, -40.0])
    R = s * np.array([[np.cos(th), -np.sin(th)], [np.sin(th), np.cos(th)]])
    b = a @ R.T + t + rng.normal(0, 0.5, a.shape)
    got = match_flow.similarity_fit(a, b)
    A = np.zeros((2 * len(a), 4))
    A[0::2] = np.c_[a[:, 0], -a[:, 1], np.ones(len(a)), np.zeros(len(a))]
    A[1::2] = np.c_[a[:, 1], a[:, 0], np.zeros(len(a)), np.ones(len(a))]
    (sa, sb, tx, ty), *_ = np.linalg.lstsq(A, b.reshape(-1), rcond=None)
    want = (np.arctan2(sb, sa), np.hypot(sa, sb), tx, ty)
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-7)
    assert abs(got[0] - th) < 1e-3 and abs(got[1] - s) < 1e-3
    assert match_flow.similarity_fit(a[:1], b[:1]) is None
    assert match_flow.similarity_fit(np.repeat(a[:1], 3, 0), b[:3]) is None


def test_pair_seeds_differ_by_pair_and_stage():
    from fieldstitch_amd import match_flow
    seeds = {match_flow.pair_seed(i, j, s) for i in range(30) for j in range(30) if i != j for s in (0x1, 0x5, 0xF)}
    assert len(seeds) == 30 * 29 * 3


def test_clahe_rejects_host_out_and_bad_images():
    """CLAHE.apply argument checks that run before any device work."""
    from fieldstitch_amd import clahe
    c = clahe.createCLAHE(2.0, (8, 8))
    assert c.getClipLimit() == 2.0 and c.getTilesGridSize() == (8, 8)
    img = np.zeros((32, 48), np.uint8)
    with pytest.raises(TypeError):
        c.apply(img, out=np.zeros_like(img))
    with pytest.raises(TypeError):
        c.apply(img.astype(np.float32))
    with pytest.raises(TypeError):
        c.apply(np.zeros((2, 32, 48), np.uint8))
