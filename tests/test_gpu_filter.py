"""Learned match filter end to end on a small flight: labels from the plain pipeline and RANSAC,
feature rows, training on the GPU kernels, threshold and evaluation."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_filter_end_to_end(gpu):
    import torch
    from fieldstitch_amd import filter_mlp as fm
    d = fm.flight_candidates(31, workers=8, nfeat=3000, rows=1, cols=3, W=1320, H=989)
    print("candidates", d["X"].shape[0], "plain", d["n_plain"], "inliers", d["n_inliers"], flush=True)
    assert d["n_inliers"] > 100 and d["n_inliers"] <= d["n_plain"]
    assert not (d["label"] & ~d["plain"]).any()          # labels are plain matches
    X = d["X"].float().cpu().numpy()
    assert np.isfinite(X).all()
    assert X[:, :128].min() >= 0 and X[:, :128].max() <= 1.0
    assert (X[:, 264:] == 0).all()
    n = (X.shape[0] // 1024) * 1024
    net = fm.MatchFilterMLP(seed=2)
    losses = fm.train(net, d["X"][:n], d["y"][:n], epochs=6, batch=1024)
    assert losses[-1] < losses[0]
    p = net.predict(d["X"]).cpu().numpy()
    tau = fm.threshold_for_recall(p, d["label"], 0.995)
    ev = fm.evaluate(p, tau, d)
    assert ev["inlier_recall"] >= 0.995
    assert 0.0 <= ev["pruned_fraction"] <= 1.0
