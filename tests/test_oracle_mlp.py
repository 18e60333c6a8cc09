"""The network oracle pinned by finite differences, torch's bfloat16 rounding and Adam's
closed form (CPU only)."""
import numpy as np

from oracle import mlp_oracle as mo


def _params(rng, scale=1.0):
    return {"W1": rng.normal(0, 0.1 * scale, (288, 256)), "b1": rng.normal(0, 0.1, 256),
            "W2": rng.normal(0, 0.1 * scale, (256, 256)), "b2": rng.normal(0, 0.1, 256),
            "w3": rng.normal(0, 0.1, 256), "b3": rng.normal(0, 0.1, 1)}


def test_bf16_matches_torch():
    import torch
    x = np.random.default_rng(0).normal(0, 10, 10000).astype(np.float32)
    x[:4] = [1.0039062, 1.01171875, -2.00390625, 3.0e-39]      # ties and a subnormal
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(mo.bf16(x), ref)


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = _params(rng)
    X = rng.normal(0, 1, (24, 288))
    y = (rng.random(24) < 0.5).astype(np.float64)
    _, g = mo.loss_and_grads(X, y, p, rounding=False)
    h = 1e-6
    for name, idx in (("W1", (3, 7)), ("W1", (250, 100)), ("b1", (5,)), ("W2", (10, 200)), ("b2", (77,)),
                      ("w3", (13,)), ("b3", (0,))):
        q = {k: np.array(v, np.float64, copy=True) for k, v in p.items()}
        q[name][idx] += h
        lp, _ = mo.loss_and_grads(X, y, q, rounding=False)
        q[name][idx] -= 2 * h
        lm, _ = mo.loss_and_grads(X, y, q, rounding=False)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[name][idx]) <= 1e-6 + 1e-4 * abs(fd), (name, idx, fd, g[name][idx])


def test_adam_closed_form():
    g = np.array([0.5, -2.0, 0.0], np.float32)
    w, m, v = mo.adam(np.ones(3), np.zeros(3), np.zeros(3), g, 1, lr=0.1)
    # first step: m_hat = g, v_hat = g^2 -> w - lr * sign(g) (eps aside)
    np.testing.assert_allclose(w, [0.9, 1.1, 1.0], rtol=1e-5)
