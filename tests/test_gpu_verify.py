"""GPU homography RANSAC equals its CPU restatement exactly; filter feature rows."""
import numpy as np
import pytest

from oracle import ransac_oracle as ro

pytestmark = pytest.mark.gpu


def _planted(rng, n, frac, H):
    a = np.c_[rng.uniform(0, 2640, n), rng.uniform(0, 1978, n)]
    w = H[2, 0] * a[:, 0] + H[2, 1] * a[:, 1] + H[2, 2]
    b = np.c_[(H[0, 0] * a[:, 0] + H[0, 1] * a[:, 1] + H[0, 2]) / w, (H[1, 0] * a[:, 0] + H[1, 1] * a[:, 1] + H[1, 2]) / w]
    b += rng.normal(0, 0.7, b.shape)
    bad = rng.random(n) >= frac
    b[bad] = np.c_[rng.uniform(0, 2640, bad.sum()), rng.uniform(0, 1978, bad.sum())]
    return a, b


def test_ransac_equals_oracle(gpu):
    from fieldstitch_amd.verify import ransac_homography
    rng = np.random.default_rng(4)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    sets = [_planted(rng, n, f, H) for n, f in ((900, 0.6), (300, 0.3), (3, 1.0), (0, 1.0), (2000, 0.9))]
    line = np.c_[np.arange(30.0), 2 * np.arange(30.0)]
    sets.append((line, line + 5.0))
    flags, Hs, cnt = ransac_homography([a for a, _ in sets], [b for _, b in sets], n_hyp=384, thr=3.0, seed=99)
    for p, (a, b) in enumerate(sets):
        rf, rh, rc = ro.ransac_pair(p, a, b, 384, 3.0, 99)
        print("pair", p, len(a), rc, flush=True)
        assert cnt[p] == rc
        np.testing.assert_array_equal(flags[p], rf)
        np.testing.assert_allclose(Hs[p].reshape(9), rh, rtol=1e-12, atol=1e-12)


def _views(rng, n, t, frac, noise=0.3):
    f, c = 1500.0, (1300.0, 980.0)
    K = np.array([[f, 0.0, c[0]], [0.0, f, c[1]], [0.0, 0.0, 1.0]])
    az = np.radians(3.0)
    R = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    X = np.c_[rng.uniform(-60, 60, n), rng.uniform(-45, 45, n), 100.0 + rng.uniform(-8, 8, n)]
    xa = (K @ X.T).T
    xb = (K @ (R @ (X - np.asarray(t, float)).T)).T
    a, b = xa[:, :2] / xa[:, 2:], xb[:, :2] / xb[:, 2:]
    b = b + rng.normal(0.0, noise, b.shape)
    bad = rng.random(n) >= frac
    b[bad] = rng.uniform(0, 2600, (int(bad.sum()), 2))
    return a, b


@pytest.mark.parametrize("model,thr", [("fundamental", 1.0), ("similarity", 4.0)])
def test_ransac_models_equal_oracle(gpu, model, thr):
    from fieldstitch_amd.verify import ransac
    rng = np.random.default_rng(12)
    sets = [_views(rng, 700, (12.0, 3.0, 1.0), 0.7), _views(rng, 250, (20.0, 2.0, 0.0), 0.5),
            _views(rng, 9, (5.0, 1.0, 0.0), 1.0), (np.zeros((0, 2)), np.zeros((0, 2))),
            _views(rng, 1500, (2.0, 9.0, 0.3), 0.9)]
    pt = np.tile([[100.0, 200.0]], (20, 1))
    sets.append((pt, pt + 5.0))
    flags, Ms, cnt = ransac(model, [a for a, _ in sets], [b for _, b in sets], n_hyp=320, thr=thr, seed=1234)
    for p, (a, b) in enumerate(sets):
        rf, rm, rc = ro.ransac_pair(p, a, b, 320, thr, 1234, model=model)
        print("pair", model, p, len(a), rc, flush=True)
        assert cnt[p] == rc
        np.testing.assert_array_equal(flags[p], rf)
        np.testing.assert_allclose(Ms[p], rm, rtol=1e-10, atol=1e-12)
    if model == "fundamental":
        assert cnt[0] > 0.5 * 700 and cnt[4] > 0.85 * 1500
