"""The RANSAC restatement pinned by planted homographies (CPU only)."""
import numpy as np

from oracle import ransac_oracle as ro


def planted(rng, n, frac, H):
    a = np.c_[rng.uniform(0, 2640, n), rng.uniform(0, 1978, n)]
    w = H[2, 0] * a[:, 0] + H[2, 1] * a[:, 1] + H[2, 2]
    b = np.c_[(H[0, 0] * a[:, 0] + H[0, 1] * a[:, 1] + H[0, 2]) / w, (H[1, 0] * a[:, 0] + H[1, 1] * a[:, 1] + H[1, 2]) / w]
    good = rng.random(n) < frac
    b[~good] = np.c_[rng.uniform(0, 2640, (~good).sum()), rng.uniform(0, 1978, (~good).sum())]
    return a, b, good


def test_recovers_planted_inliers():
    rng = np.random.default_rng(0)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    a, b, good = planted(rng, 800, 0.6, H)
    flags, h, cnt = ro.ransac_pair(3, a, b, 256, 3.0, 7)
    np.testing.assert_array_equal(flags, good)
    assert cnt == good.sum()
    np.testing.assert_allclose(h.reshape(3, 3) / h[8], H / H[2, 2], rtol=1e-6, atol=1e-6)


def test_degenerate_and_small_sets():
    rng = np.random.default_rng(1)
    flags, h, cnt = ro.ransac_pair(0, np.zeros((3, 2)), np.zeros((3, 2)), 64, 3.0, 1)
    assert cnt == 0 and not flags.any()
    # all points on a line: every 4-point sample is degenerate
    a = np.c_[np.arange(20.0), 2 * np.arange(20.0)]
    flags, h, cnt = ro.ransac_pair(1, a, a + 5.0, 64, 3.0, 1)
    assert cnt == 0 and not flags.any()


def test_draws_distinct():
    idx, ok = ro.samples(5, 2, 6, np.arange(500))
    assert ok.all()
    assert all(len(set(r)) == 4 for r in idx)
