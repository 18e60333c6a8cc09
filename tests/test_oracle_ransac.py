"""The RANSAC restatement pinned by planted homographies, two-view geometry and similarities (CPU only)."""
import numpy as np
import pytest

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


def _two_views(rng, n, t, rot_deg=(2.0, -1.5, 4.0), f=1500.0, c=(1300.0, 980.0)):
    """Points on a bumpy ground 100 m below camera a, camera b displaced by t and rotated:
    pixel correspondences and the true F (x_b^T F x_a = 0)."""
    K = np.array([[f, 0.0, c[0]], [0.0, f, c[1]], [0.0, 0.0, 1.0]])
    ax, ay, az = np.radians(rot_deg)
    Rx = np.array([[1, 0, 0], [0, np.cos(ax), -np.sin(ax)], [0, np.sin(ax), np.cos(ax)]])
    Ry = np.array([[np.cos(ay), 0, np.sin(ay)], [0, 1, 0], [-np.sin(ay), 0, np.cos(ay)]])
    Rz = np.array([[np.cos(az), -np.sin(az), 0], [np.sin(az), np.cos(az), 0], [0, 0, 1]])
    R = Rz @ Ry @ Rx
    X = np.c_[rng.uniform(-60, 60, n), rng.uniform(-45, 45, n), 100.0 + rng.uniform(-8, 8, n)]
    xa = (K @ X.T).T
    xa = xa[:, :2] / xa[:, 2:]
    Xb = (R @ (X - np.asarray(t, float)).T).T
    xb = (K @ Xb.T).T
    xb = xb[:, :2] / xb[:, 2:]
    tx = np.array([[0, -t[2], t[1]], [t[2], 0, -t[0]], [-t[1], t[0], 0]], float)
    # x_b ~ K R (X - t), x_a ~ K X  =>  E = R [t]x with X_b = R (X - t)
    E = R @ tx
    Kinv = np.linalg.inv(K)
    F = Kinv.T @ E @ Kinv
    return xa, xb, F / np.linalg.norm(F)


def _epi(F, a, b):
    F = F.reshape(3, 3)
    ha = np.c_[a, np.ones(len(a))]
    hb = np.c_[b, np.ones(len(b))]
    l2 = ha @ F.T
    return np.abs(np.sum(hb * l2, 1)) / np.hypot(l2[:, 0], l2[:, 1])


@pytest.mark.parametrize("t", [(3.0, 1.0, 0.5), (20.0, 2.0, 0.0)])
def test_fundamental_8point_exact(t):
    """Exact correspondences: the 8-point solve reproduces the true F up to scale, also for the
    nadir sideways translation whose normalized F has a near-zero (3, 3) entry."""
    rng = np.random.default_rng(5)
    a, b, Ft = _two_views(rng, 40, t)
    F, ok = ro.solve_f8(a[None, :8, 0], a[None, :8, 1], b[None, :8, 0], b[None, :8, 1])
    assert ok[0]
    Fn = F[0] / np.linalg.norm(F[0])
    Fn = Fn * np.sign(Fn @ Ft.ravel())
    assert np.abs(Fn - Ft.ravel()).max() < 1e-6
    assert _epi(F[0], a, b).max() < 1e-6                          # every other point too
    # a transposed or sign-flipped model does not fit
    assert _epi(F[0].reshape(3, 3).T.ravel(), a, b).max() > 1e-2


def test_fundamental_ransac_with_outliers():
    rng = np.random.default_rng(6)
    a, b, Ft = _two_views(rng, 300, (12.0, 3.0, 1.0))
    b = b + rng.normal(0.0, 0.2, b.shape)
    out = rng.random(300) < 0.3
    b[out] = rng.uniform(0, 2600, (int(out.sum()), 2))
    flags, F, cnt = ro.ransac_pair(0, a, b, 1024, 1.0, 77, model="fundamental")
    true_in = ~out
    assert cnt == flags.sum()
    assert (flags & true_in).sum() >= 0.95 * true_in.sum()
    assert (flags & out).sum() <= 0.05 * out.sum()
    assert np.median(_epi(F, a[true_in], b[true_in])) < 0.5


def test_similarity_ransac_planted():
    rng = np.random.default_rng(7)
    n = 200
    a = rng.uniform(0, 2000, (n, 2))
    ang, sc, tx, ty = 0.3, 1.07, 40.0, -25.0
    A = sc * np.array([[np.cos(ang), -np.sin(ang)], [np.sin(ang), np.cos(ang)]])
    b = a @ A.T + [tx, ty] + rng.normal(0.0, 0.5, (n, 2))
    out = rng.random(n) < 0.25
    b[out] = rng.uniform(0, 2000, (int(out.sum()), 2))
    flags, M, cnt = ro.ransac_pair(3, a, b, 256, 3.0, 9, model="similarity")
    assert (flags == ~out).mean() > 0.97
    assert abs(np.arctan2(M[3], M[0]) - ang) < 0.01 and abs(np.hypot(M[0], M[3]) - sc) < 0.01
    assert M[1] == -M[3] and M[4] == M[0]


def test_fundamental_and_similarity_degenerate():
    a = np.tile([[100.0, 200.0]], (20, 1))
    b = a + 5.0
    for model in ("fundamental", "similarity"):
        flags, M, cnt = ro.ransac_pair(0, a, b, 64, 1.0, 1, model=model)
        assert cnt == 0 and not flags.any() and not M.any()


def test_rank2_projection_is_the_svd_truncation():
    """The Jacobi projection equals the textbook closest rank-2 matrix (numpy's SVD with the
    smallest singular value zeroed) of the scaled matrix, and leaves a rank-2 matrix as it is."""
    rng = np.random.default_rng(11)
    f = rng.normal(size=(200, 9))
    f[:20, 6:] = f[:20, :3] * 2.0 - f[:20, 3:6]              # already rank 2
    f[20:40] *= np.array([1.0, 1e-3, 1.0, 1e-6, 1.0, 1.0, 1e3, 1.0, 1.0])    # badly scaled
    got = ro.rank2_project(f)
    for k in range(len(f)):
        m = (f[k] / np.abs(f[k]).max()).reshape(3, 3)
        U, s, Vt = np.linalg.svd(m)
        want = (U[:, :2] * s[:2]) @ Vt[:2]
        np.testing.assert_allclose(got[k].reshape(3, 3), want, atol=1e-12 * s[0])
        assert np.linalg.svd(got[k].reshape(3, 3), compute_uv=False)[2] <= 1e-12 * s[0]
    np.testing.assert_allclose(got[:20], f[:20] / np.abs(f[:20]).max(1, keepdims=True), atol=1e-12)
    assert not ro.rank2_project(np.zeros((1, 9))).any()


def test_fundamental_ransac_model_is_rank2_and_labels_its_inliers():
    """Planted two-view case: the winning F is rank 2, the flags and count are exactly the
    inliers of that rank-2 model, and the true inliers are found."""
    rng = np.random.default_rng(8)
    a, b, Ft = _two_views(rng, 400, (15.0, 4.0, 0.5))
    b = b + rng.normal(0.0, 0.3, b.shape)
    out = rng.random(400) < 0.25
    b[out] = rng.uniform(0, 2600, (int(out.sum()), 2))
    flags, F, cnt = ro.ransac_pair(2, a, b, 1024, 1.0, 5, model="fundamental")
    s = np.linalg.svd(F.reshape(3, 3), compute_uv=False)
    assert s[2] <= 1e-12 * s[0]
    np.testing.assert_array_equal(flags, ro.inliers_f(F[None], a, b, 1.0)[0])
    assert cnt == flags.sum() and (flags & ~out).sum() >= 0.95 * (~out).sum() and (flags & out).sum() <= 3


def _dlt_lstsq(a, b):
    """The same normalized least-squares DLT with h33 = 1, solved by numpy's lstsq on the
    stacked rows (not the normal equations)."""
    ca, cb = a.mean(0), b.mean(0)
    sa = np.sqrt(2.0) / np.linalg.norm(a - ca, axis=1).mean()
    sb = np.sqrt(2.0) / np.linalg.norm(b - cb, axis=1).mean()
    x, y = ((a - ca) * sa).T
    u, v = ((b - cb) * sb).T
    z, o = np.zeros(len(a)), np.ones(len(a))
    M = np.r_[np.c_[x, y, o, z, z, z, -u * x, -u * y], np.c_[z, z, z, x, y, o, -v * x, -v * y]]
    h = np.r_[np.linalg.lstsq(M, np.r_[u, v], rcond=None)[0], 1.0].reshape(3, 3)
    Ta = np.array([[sa, 0, -sa * ca[0]], [0, sa, -sa * ca[1]], [0, 0, 1.0]])
    Tb = np.array([[sb, 0, -sb * cb[0]], [0, sb, -sb * cb[1]], [0, 0, 1.0]])
    H = np.linalg.inv(Tb) @ h @ Ta
    return H / H[2, 2]


def _reproj(H, a, b):
    H = H.reshape(3, 3)
    p = np.c_[a, np.ones(len(a))] @ H.T
    return np.sqrt(np.mean(np.sum((p[:, :2] / p[:, 2:] - b) ** 2, 1)))


def test_homography_refit_on_inliers():
    """The winning homography is refit on its inliers: it equals an independent least-squares
    solve of the same normalized system, fits the true inliers better than the 4-point winner,
    and the flags / count stay the winner's."""
    rng = np.random.default_rng(21)
    H = np.array([[1.01, -0.04, 550.0], [0.03, 0.98, -60.0], [2e-6, -1e-6, 1.0]])
    a, b, good = planted(rng, 900, 0.55, H)
    b[good] += rng.normal(0.0, 0.8, (int(good.sum()), 2))
    flags, h, cnt = ro.ransac_pair(4, a, b, 256, 3.0, 13)
    K, solve, test = ro.MODELS["homography"]
    idx, ok = ro.samples(13, 4, len(a), np.arange(256), K)
    hyp, valid = solve(a[idx, 0], a[idx, 1], b[idx, 0], b[idx, 1])
    counts = np.where(valid & ok, test(hyp, a, b, 9.0).sum(1), -1)
    win = hyp[int(np.argmax(counts))]
    np.testing.assert_array_equal(flags, test(win[None], a, b, 9.0)[0])
    assert cnt == counts.max() == flags.sum()
    np.testing.assert_allclose(h.reshape(3, 3), _dlt_lstsq(a[flags], b[flags]), rtol=1e-9, atol=1e-9)
    assert _reproj(h, a[good], b[good]) < _reproj(win, a[good], b[good])
    assert _reproj(h, a[good], b[good]) < 1.2
    # fewer than 5 inliers: the hypothesis' model is kept
    few = np.zeros(len(a), bool)
    few[np.nonzero(flags)[0][:4]] = True
    assert ro.refit_h(a, b, few, win) is win


def test_fma_is_correctly_rounded():
    """The oracle's fused multiply-add (C99 fma) equals a * b + c computed exactly in rationals
    and rounded once; the unfused expression differs on the cases built to show it."""
    from fractions import Fraction
    from oracle.sift_oracle import fma
    rng = np.random.default_rng(31)
    a = rng.normal(size=2000) * 10.0 ** rng.integers(-6, 7, 2000)
    b = rng.normal(size=2000) * 10.0 ** rng.integers(-6, 7, 2000)
    c = -(a * b) * (1.0 + rng.normal(size=2000) * 1e-12)     # heavy cancellation
    e = 2.0 ** -52
    a = np.r_[a, 1.0 + e, 3.0, 0.1]
    b = np.r_[b, 1.0 - e, 1.0 / 3.0, 10.0]
    c = np.r_[c, -1.0, -1.0, -1.0]
    got = fma(a, b, c)
    want = np.array([float(Fraction(x) * Fraction(y) + Fraction(z)) for x, y, z in zip(a, b, c)])
    np.testing.assert_array_equal(got, want)
    assert got[-3] == -e * e and (a[-3] * b[-3] + c[-3]) == 0.0
    assert np.count_nonzero(got != a * b + c) > 1000
    # broadcasting as numpy's operators
    assert fma(np.ones((3, 1)), np.arange(4.0)[None, :], 1.0).shape == (3, 4)


def test_inlier_tests_equal_the_geometric_errors():
    """Each model's fused inlier test agrees with its geometric error computed the long way
    (homography: |H x_a / w - x_b|^2; fundamental: the larger squared distance of x_b to the
    line F x_a and of x_a to the line F^T x_b; similarity: |M x_a - x_b|^2) against thr^2,
    wherever the error is not within 1e-9 relative of the threshold."""
    rng = np.random.default_rng(32)
    a = rng.uniform(0, 2640, (3000, 2))
    H = np.array([[1.01, -0.04, 550.0], [0.03, 0.98, -60.0], [2e-6, -1e-6, 1.0]])
    w = a @ H[2, :2] + H[2, 2]
    b = np.c_[(a @ H[0, :2] + H[0, 2]) / w, (a @ H[1, :2] + H[1, 2]) / w] + rng.normal(0, 3.0, a.shape)
    ha, hb = np.c_[a, np.ones(len(a))], np.c_[b, np.ones(len(b))]
    p = ha @ H.T
    err_h = np.sum((p[:, :2] / p[:, 2:] - b) ** 2, 1)
    F = rng.normal(size=(3, 3))
    l2, l1 = ha @ F.T, hb @ F
    err_f = np.maximum(np.sum(hb * l2, 1) ** 2 / (l2[:, 0] ** 2 + l2[:, 1] ** 2),
                       np.sum(ha * l1, 1) ** 2 / (l1[:, 0] ** 2 + l1[:, 1] ** 2))
    M = np.array([1.02, -0.05, 30.0, 0.05, 1.02, -12.0, 0.0, 0.0, 1.0])
    err_s = np.sum((a @ M.reshape(3, 3)[:2, :2].T + M.reshape(3, 3)[:2, 2] - b) ** 2, 1)
    for name, model, err in (("homography", H.ravel(), err_h), ("fundamental", F.ravel(), err_f),
                             ("similarity", M, err_s)):
        thr2 = float(np.median(err))
        got = ro.MODELS[name][2](model[None], a, b, thr2)[0]
        clear = np.abs(err - thr2) > 1e-9 * thr2
        assert 0.3 < got.mean() < 0.7, name
        np.testing.assert_array_equal(got[clear], (err < thr2)[clear], err_msg=name)
    # a transposed F is a different test
    got_t = ro.inliers_f(F.T.ravel()[None], a, b, float(np.median(err_f)))[0]
    assert (got_t != (err_f < np.median(err_f))).mean() > 0.1


def test_fundamental_zero_normal_never_passes():
    """A point whose epipolar line has a zero normal (here F = 0, and F with a zero first two
    columns for every point) is no inlier, as under cv2's infinite error."""
    a = np.random.default_rng(33).uniform(0, 100, (50, 2))
    assert not ro.inliers_f(np.zeros((1, 9)), a, a, 1e30).any()
    F = np.zeros((1, 9))
    F[0, [2, 5, 8]] = 1.0                      # l2 = (1, 1, 1) for every point, l1 = (0, 0, .)
    assert not ro.inliers_f(F, a, a, 1e30).any()


# ---- adaptive stopping (fsx_ransac_batch_adaptive) ---------------------------------------------

def test_stopping_rule_is_the_closed_form():
    """confident(c, n, K, R, 1 - conf) holds exactly when 256 R >= log(1 - conf) / log(1 - w^K)
    (the iteration count cv2's estimators stop at), away from the boundary; a power computed
    with one squaring too few or too many, or w^K with a wrong exponent, fails it."""
    import math
    checked = 0
    for K in (2, 4, 8):
        for conf in (0.99, 0.995, 0.999):
            for n in (37, 500, 2612):
                for c in range(1, n + 1, max(1, n // 97)):
                    w = c / n
                    denom = math.log1p(-(w ** K)) if w < 1 else -math.inf
                    need = math.log(1.0 - conf) / denom if denom != 0 else math.inf
                    for R in (1, 2, 3, 5, 8, 16, 39, 40):
                        if math.isfinite(need) and abs(256 * R - need) <= 1e-9 * need:
                            continue
                        assert ro.confident(c, n, K, R, 1.0 - conf) == (256 * R >= need), (K, conf, n, c, R)
                        checked += 1
    assert checked > 10000


def test_unreachable_confidence_equals_the_fixed_count():
    """With 1 - conf = 2^-52 and few inliers the rule never holds within the cap, so the
    adaptive form draws every hypothesis and returns what the fixed-count RANSAC returns."""
    rng = np.random.default_rng(31)
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    a, b, _ = planted(rng, 400, 0.3, H)
    conf = 1.0 - 2.0 ** -52
    for model, cap, thr in (("homography", 700, 3.0), ("fundamental", 600, 1.0)):
        f1, m1, c1 = ro.ransac_pair(0, a, b, cap, thr, 77, model=model)
        f2, m2, c2, used = ro.ransac_pair_adaptive(0, a, b, cap, conf, thr, 77, model=model)
        assert used == cap and c1 == c2
        np.testing.assert_array_equal(f1, f2)
        np.testing.assert_array_equal(m1, m2)


def test_adaptive_stops_at_the_first_confident_round():
    """All inliers: the first round's best model has w = 1, so one round of 256; 70 % inliers:
    the pair stops once 256 R reaches the closed-form count for its final inlier ratio, far
    below the cap, with the planted inliers found; fewer than K matches: one round, no model."""
    import math
    rng = np.random.default_rng(32)
    n = 300
    a = rng.uniform(0, 2000, (n, 2))
    b = a * 1.02 + [15.0, -7.0]
    f, M, c, used = ro.ransac_pair_adaptive(0, a, b, 2000, 0.99, 3.0, 5, model="similarity")
    assert used == 256 and c == n and f.all()
    H = np.array([[0.99, -0.05, 600.0], [0.05, 0.99, -40.0], [1e-6, -2e-6, 1.0]])
    a, b, good = planted(rng, 900, 0.7, H)
    f, h, c, used = ro.ransac_pair_adaptive(0, a, b, 2000, 0.995, 3.0, 11)
    need = math.log(0.005) / math.log1p(-((c / 900) ** 4))
    assert used % 256 == 0 and 256 <= used < 2000 and used >= need
    np.testing.assert_array_equal(f, good)
    f, h, c, used = ro.ransac_pair_adaptive(0, a[:3], b[:3], 2000, 0.995, 3.0, 11)
    assert used == 256 and c == 0 and not f.any()
