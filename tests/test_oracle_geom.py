"""The geometry oracle pinned by closed forms, round trips and scipy (CPU only)."""
import numpy as np
import pytest

from oracle import geom_oracle as go

K_MILD = (3000.0, 1320.0, 989.0, -0.08, 0.02, 3e-4, -2e-4, -0.003)


def test_rescale_closed_form():
    xy = np.array([[0, 0], [10.5, 3.25], [659.0, 493.0]], np.float32)
    out = go.rescale(xy, 2640, 1978, 660, 494)
    x64 = xy.astype(np.float64)
    exp = np.c_[(x64[:, 0] + 0.5) * 4.0 - 0.5, (x64[:, 1] + 0.5) * (1978 / 494.0) - 0.5]
    np.testing.assert_allclose(out, exp, rtol=1e-15)
    np.testing.assert_array_equal(go.rescale(xy, 660, 494, 660, 494), xy.astype(np.float64))


def test_undistort_zero_coefficients_is_affine_round_trip():
    rng = np.random.default_rng(0)
    uv = rng.uniform(0, 3000, (500, 2))
    k = (2500.0, 1500.0, 1000.0, 0, 0, 0, 0, 0)
    out = go.undistort_pixels(uv, k)
    exp = np.c_[((uv[:, 0] - 1500.0) / 2500.0) * 2500.0 + 1500.0, ((uv[:, 1] - 1000.0) / 2500.0) * 2500.0 + 1000.0]
    np.testing.assert_array_equal(out, exp)


@pytest.mark.parametrize("k", [K_MILD, (3000.0, 1320.0, 989.0, 0.05, 0, 0, 0, 0),
                               (3000.0, 1320.0, 989.0, 0, 0, 0, 5e-4, 0), (3000.0, 1320.0, 989.0, 0, 0, -5e-4, 0, 0)])
def test_undistort_inverts_forward_model(k):
    # every coefficient enters the forward model; a dropped or sign-flipped term breaks the trip
    rng = np.random.default_rng(1)
    ideal = rng.uniform([0, 0], [2640, 1978], (2000, 2))
    dist = go.distort_pixels(ideal, k)
    back = go.undistort_pixels(dist, k)
    assert np.abs(back - ideal).max() < 1e-6


def test_undistort_radius_clamp():
    # strongly negative k1 drives rad below 0.2 for far points: the clamp keeps it finite
    k = (1000.0, 0.0, 0.0, -2.0, 0, 0, 0, 0)
    out = go.undistort_pixels(np.array([[900.0, 900.0]]), k)
    assert np.isfinite(out).all()


def _cam(yaw_deg, C, f=3000.0, cx=1320.0, cy=989.0):
    a = np.radians(yaw_deg)
    # nadir camera: camera z looks down (-world z), x along the yawed ground axis
    Rz = np.array([[np.cos(a), np.sin(a), 0], [-np.sin(a), np.cos(a), 0], [0, 0, 1]])
    flip = np.diag([1.0, -1.0, -1.0])
    R = flip @ Rz
    K = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])
    return K, R, np.asarray(C, np.float64)


def test_ground_predict_recovers_projection():
    Ki, Ri, Ci = _cam(10.0, [0.0, 0.0, 120.0])
    Kj, Rj, Cj = _cam(-5.0, [30.0, 8.0, 118.0])
    rng = np.random.default_rng(2)
    G = np.c_[rng.uniform(-30, 60, 300), rng.uniform(-25, 30, 300), np.full(300, 2.5)]
    ui, zi = go.project(G, Ki, Ri, Ci)
    uj, zj = go.project(G, Kj, Rj, Cj)
    assert (zi > 0).all() and (zj > 0).all()
    pred, ok = go.ground_predict(ui, np.linalg.inv(Ki), Ri, Ci, 2.5, Rj, Kj, Cj)
    assert ok.all()
    assert np.abs(pred - uj).max() < 1e-6


def test_ground_predict_behind_camera():
    Ki, Ri, Ci = _cam(0.0, [0.0, 0.0, 120.0])
    Kj, Rj, Cj = _cam(0.0, [10.0, 0.0, 120.0])
    # ground above the camera: t < 0
    pred, ok = go.ground_predict(np.array([[1320.0, 989.0]]), np.linalg.inv(Ki), Ri, Ci, 200.0, Rj, Kj, Cj)
    assert not ok[0]


def _grid(rows, cols, step, jitter, seed):
    rng = np.random.default_rng(seed)
    g = np.array([(c * step, r * step * 1.3) for r in range(rows) for c in range(cols)], np.float64)
    return g + rng.normal(0, jitter, g.shape)


@pytest.mark.parametrize("n_rows,n_cols,neigh,has_rel", [(5, 6, 10, True), (3, 9, 4, False), (1, 2, 16, True)])
def test_pairs_brute_equals_kdtree(n_rows, n_cols, neigh, has_rel):
    xy = _grid(n_rows, n_cols, 30.0, 2.0, n_rows * 10 + n_cols)
    hl = np.full(len(xy), 40.0)
    assert go.select_pairs_brute(xy, hl, neigh, has_rel) == go.select_pairs_kdtree(xy, hl, neigh, has_rel)


def test_pairs_hand_case():
    # 4 photos on a line 10 m apart, half_long 6: the 0.95 limit admits 10 m (< 11.4), not 20 m
    xy = np.array([[0, 0], [10, 0], [20, 0], [30, 0]], np.float64)
    got = go.select_pairs_brute(xy, np.full(4, 6.0), 2, True)
    assert got == [(0, 1), (1, 2), (2, 3)]
    # without a relative altitude every kNN neighbour counts; temporal +2 needs < 12 m
    got = go.select_pairs_brute(xy, np.full(4, 6.0), 2, False)
    assert got == [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]
