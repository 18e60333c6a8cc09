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


def test_filter_gain_on_held_out_flight_and_pruning_stage(gpu):
    """Trained on two small flights, threshold from a third, evaluated on a fourth: at >= 99.5 %
    of the verified inliers kept, the outlier share reaching verification falls at least 5x and
    some candidates are pruned.  Then the pruning stage inside the matcher (run_filtered) gives
    exactly the one-query-per-target rule applied to the plain run's accepted queries that
    score >= tau, and with no network it gives the plain run's lists."""
    import torch
    from fieldstitch_amd import filter_mlp as fm
    from fieldstitch_amd import matching
    kw = dict(nfeat=3000, rows=2, cols=4, W=1320, H=989, workers=1)
    train = [fm.flight_candidates(s, **kw) for s in (41, 42)]
    val = fm.flight_candidates(43, **kw)
    test = fm.flight_candidates(44, keep_matcher=True, **kw)
    X = torch.cat([d["X"] for d in train])
    y = torch.cat([d["y"] for d in train])
    n = (X.shape[0] // 1024) * 1024
    net = fm.MatchFilterMLP(seed=2)
    losses = fm.train(net, X[:n], y[:n], epochs=8, batch=1024)
    tau = fm.threshold_for_recall(net.predict(val["X"]).cpu().numpy(), val["label"], 0.997)
    ev = fm.evaluate(net.predict(test["X"]).cpu().numpy(), tau, test)
    out_plain = 1.0 - ev["plain_ratio_test"]["inlier_fraction"]
    out_filt = 1.0 - ev["filter_then_ratio_test"]["inlier_fraction"]
    print("losses", [round(v, 4) for v in losses], "tau", tau, "outlier share", out_plain, "->", out_filt,
          "recall", ev["filter_then_ratio_test"]["inlier_recall"], "pruned", ev["pruned_fraction"], flush=True)
    assert ev["filter_then_ratio_test"]["inlier_recall"] >= 0.995
    assert ev["pruned_fraction"] > 0.0
    assert out_plain > 0.0 and out_filt * 5.0 <= out_plain
    # the pruning stage of the matcher
    pm, sizes, angles, qoff = test["matcher"], test["sizes"], test["angles"], test["query_offsets"]
    pm.run("mfma")
    plain = pm.results()
    best, b1, _ = matching._query_state(pm)
    p_all = net.predict(pm.filter_features(sizes, angles)).cpu().numpy()
    pm.run_filtered("mfma", None, tau, None, None)
    for a, b in zip(pm.results(), plain):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    pm.run_filtered("mfma", net, tau, sizes, angles)
    got = pm.results()
    scored, kept = pm.filter_counts()
    assert scored >= sum(len(q) for q, _ in plain)
    n_in, n_in_kept = 0, 0
    for p, (g_q, g_t) in enumerate(got):
        lo = qoff[p]
        hi = qoff[p + 1] if p + 1 < len(qoff) else len(best)
        q = np.arange(hi - lo)
        ok = (best[lo:hi] >= 0) & (p_all[lo:hi] >= tau)
        wq, wt = matching._unique(q[ok], best[lo:hi][ok].astype(np.int64), b1[lo:hi][ok])
        assert np.array_equal(g_q, wq) and np.array_equal(g_t, wt), p
        lab = test["query_label"][lo:hi]
        n_in += int(lab.sum())
        n_in_kept += int(lab[g_q].sum())
    assert kept < scored and n_in_kept >= 0.995 * n_in
