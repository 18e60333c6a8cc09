"""GPU geometry (rescale + undistortion, ground-plane prior, pair selection) against the
oracle, through the C ABI."""
import os

import numpy as np
import pytest

from oracle import geom_oracle as go

pytestmark = pytest.mark.gpu

KS = [(3000.0, 1320.0, 989.0, 0, 0, 0, 0, 0),
      (3000.0, 1320.0, 989.0, -0.08, 0.02, 3e-4, -2e-4, -0.003),
      (1800.0, 1000.0, 700.0, 0.12, -0.05, -1e-3, 8e-4, 0.01),
      (1000.0, 0.0, 0.0, -2.0, 0, 0, 0, 0)]          # drives |rad| under the 0.2 clamp


def test_undistort_batch_bitexact(gpu):
    import torch
    from fieldstitch_amd import camera
    rng = np.random.default_rng(5)
    items, host = [], []
    for s, (k, n, (W, H), (rw, rh)) in enumerate(zip(KS + [KS[1]], [3000, 1, 777, 64, 0],
                                                     [(2640, 1978)] * 5, [(660, 494), (2640, 1978), (1320, 989),
                                                                          (660, 494), (660, 494)])):
        xy = rng.uniform([0, 0], [rw, rh], (n, 2)).astype(np.float32)
        print("set", s, "k", k, "n", n, flush=True)
        items.append((torch.from_numpy(xy).cuda(), (W / float(rw), H / float(rh)), k))
        host.append((xy, W, H, rw, rh, k))
    out = camera.undistort_batch(items)
    for (xy, W, H, rw, rh, k), (d_xy, d_und) in zip(host, out):
        ref_xy = go.rescale(xy, W, H, rw, rh)
        np.testing.assert_array_equal(d_xy.cpu().numpy(), ref_xy.reshape(-1, 2))
        np.testing.assert_array_equal(d_und.cpu().numpy(), go.undistort_pixels(ref_xy, k).reshape(-1, 2))


@pytest.mark.parametrize("k", KS)
def test_undistort_pixels_f64_bitexact(gpu, k):
    from fieldstitch_amd import camera
    uv = np.random.default_rng(6).uniform(-100, 2800, (4096, 2))
    np.testing.assert_array_equal(camera.undistort_pixels(uv, k), go.undistort_pixels(uv, k))


def _cam(yaw_deg, C, f=3000.0, cx=1320.0, cy=989.0):
    a = np.radians(yaw_deg)
    Rz = np.array([[np.cos(a), np.sin(a), 0], [-np.sin(a), np.cos(a), 0], [0, 0, 1]])
    R = np.diag([1.0, -1.0, -1.0]) @ Rz
    K = np.array([[f, 0, cx], [0, f, cy], [0, 0, 1.0]])
    return K, R, np.asarray(C, np.float64)


def test_ground_predict_vs_oracle(gpu):
    from fieldstitch_amd.matching import ground_predict
    Ki, Ri, Ci = _cam(10.0, [0.0, 0.0, 120.0])
    Kj, Rj, Cj = _cam(-5.0, [30.0, 8.0, 118.0])
    rng = np.random.default_rng(7)
    G = np.c_[rng.uniform(-30, 60, 5000), rng.uniform(-25, 30, 5000), np.full(5000, 2.5)]
    ui, _ = go.project(G, Ki, Ri, Ci)
    uj, _ = go.project(G, Kj, Rj, Cj)
    args = (np.linalg.inv(Ki), Ri, Ci, 2.5, Rj, Kj, Cj)
    pred, ok = ground_predict(ui, *args)
    ref, ref_ok = go.ground_predict(ui, *args)
    np.testing.assert_array_equal(ok, ref_ok)
    # numpy's matmul goes through BLAS (order / FMA not fixed): float64 ulps, not bits
    np.testing.assert_allclose(pred, ref, rtol=1e-12, atol=1e-9)
    assert np.abs(pred - uj).max() < 1e-6
    # a ground plane above the camera gives t < 0
    pred, ok = ground_predict(ui[:10], np.linalg.inv(Ki), Ri, Ci, 500.0, Rj, Kj, Cj)
    assert not ok.any()


def _grid(rows, cols, step, jitter, seed):
    rng = np.random.default_rng(seed)
    g = np.array([(c * step, r * step * 1.3) for r in range(rows) for c in range(cols)], np.float64)
    return g + rng.normal(0, jitter, g.shape)


@pytest.mark.parametrize("rows,cols,neigh,has_rel", [(20, 20, 16, True), (5, 6, 10, True), (3, 9, 4, False),
                                                     (1, 2, 16, True), (7, 11, 64, True)])
def test_select_pairs_vs_oracle(gpu, rows, cols, neigh, has_rel):
    from fieldstitch_amd import pairs
    xy = _grid(rows, cols, 30.0, 2.0, rows * 100 + cols)
    hl = np.random.default_rng(rows).uniform(30, 45, len(xy))
    print("grid", rows, cols, neigh, has_rel, flush=True)
    assert pairs.select_pairs(xy, hl, neigh, has_rel) == go.select_pairs_kdtree(xy, hl, neigh, has_rel)
    assert pairs.select_pairs(xy, hl, neigh, has_rel) == go.select_pairs_brute(xy, hl, neigh, has_rel)


def test_select_pairs_hand_case(gpu):
    from fieldstitch_amd import pairs
    xy = np.array([[0, 0], [10, 0], [20, 0], [30, 0]], np.float64)
    assert pairs.select_pairs(xy, np.full(4, 6.0), 2, True) == [(0, 1), (1, 2), (2, 3)]
    assert pairs.select_pairs(xy, np.full(4, 6.0), 2, False) == [(0, 1), (0, 2), (1, 2), (1, 3), (2, 3)]
    assert pairs.select_pairs(xy[:1], np.full(1, 6.0), 2, True) == []


def test_get_and_predict_from_context(gpu, tmp_path):
    from fieldstitch_amd import matching
    rng = np.random.default_rng(8)
    W, H, rw, rh = 2640, 1978, 1320, 989
    ks = [KS[1], KS[2]]
    for i in range(2):
        xy = rng.uniform([0, 0], [rw, rh], (300, 2)).astype(np.float32)
        desc = rng.integers(0, 256, (300, 128), dtype=np.uint8)
        np.savez(os.path.join(tmp_path, "%d.npz" % i), xy=xy, desc=desc, resp=np.ones(300, np.float32),
                 red_wh=np.array([rw, rh], np.int32))
    Ks, Rs, Cs = [], [], []
    for i, (yaw, C) in enumerate([(3.0, [0.0, 0.0, 110.0]), (-2.0, [25.0, 5.0, 112.0])]):
        K, R, Cc = _cam(yaw, C, f=ks[i][0], cx=ks[i][1], cy=ks[i][2])
        Ks.append(K), Rs.append(R), Cs.append(Cc)
    ctx = {"feat_dir": str(tmp_path), "W": [W, W], "H": [H, H], "K": Ks,
           "dist": [np.array(k[3:]) for k in ks], "R": Rs, "C": Cs, "zg": [1.0, 1.0]}
    matching.init_worker(ctx)
    A = matching._get(0)
    raw = np.load(os.path.join(tmp_path, "0.npz"))
    ref_xy = go.rescale(raw["xy"], W, H, rw, rh)
    np.testing.assert_array_equal(A["xy"], ref_xy)
    np.testing.assert_array_equal(A["und"], go.undistort_pixels(ref_xy, ks[0]))
    np.testing.assert_array_equal(A["desc"], raw["desc"].astype(np.float32))
    assert matching._get(0) is A                                    # cached
    xy, desc = matching.load_features(str(tmp_path), 1, W, H)
    np.testing.assert_array_equal(xy, go.rescale(np.load(os.path.join(tmp_path, "1.npz"))["xy"], W, H, rw, rh))
    pred, ok = matching._predict(A["und"], 0, 1)
    ref, ref_ok = go.ground_predict(A["und"], np.linalg.inv(Ks[0]), Rs[0], Cs[0], 1.0, Rs[1], Ks[1], Cs[1])
    np.testing.assert_array_equal(ok, ref_ok)
    np.testing.assert_allclose(pred, ref, rtol=1e-12, atol=1e-9)
