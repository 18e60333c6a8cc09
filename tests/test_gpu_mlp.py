"""Learned match filter kernels (GEMM variants, head, column sums, Adam) against the numpy
oracle, and a training run of the network on the GPU."""
import ctypes as C

import numpy as np
import pytest

from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu


def _bf(t, rng, shape, scale=1.0):
    import torch
    a = mo.bf16(rng.normal(0, scale, shape).astype(np.float32))
    return torch.from_numpy(a).to(torch.bfloat16).cuda(), a.astype(np.float64)


@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("tb", [0, 1])
def test_gemm_variants(gpu, ta, tb):
    import torch
    from fieldstitch_amd import _lib
    rng = np.random.default_rng(10 + 2 * ta + tb)
    M, N, K = 100, 70, 96
    A, a = _bf(torch, rng, (K, M) if ta else (M, K))
    Bm, b = _bf(torch, rng, (N, K) if tb else (K, N))
    opa = a.T if ta else a
    opb = b.T if tb else b
    ref = opa @ opb
    lda = M if ta else K
    ldb = K if tb else N
    P = lambda t: C.c_void_p(t.data_ptr())   # noqa: E731
    st = C.c_void_p(_lib.stream_handle())
    c32 = torch.empty((M, N), dtype=torch.float32, device="cuda")
    _lib.call("fsx_mlp_gemm", ta, tb, 1, P(A), lda, P(Bm), ldb, M, N, K, P(c32), None, N, None, None, st)
    np.testing.assert_allclose(c32.cpu().numpy(), ref, rtol=1e-5, atol=1e-4)
    bias = torch.from_numpy(rng.normal(0, 1, N).astype(np.float32)).cuda()
    c16 = torch.empty((M, N), dtype=torch.bfloat16, device="cuda")
    _lib.call("fsx_mlp_gemm", ta, tb, 0, P(A), lda, P(Bm), ldb, M, N, K, None, P(c16), N, P(bias), None, st)
    exp = mo.bf16(np.maximum(ref + bias.cpu().numpy()[None, :], 0).astype(np.float32))
    got = c16.float().cpu().numpy()
    assert np.abs(got - exp).max() <= 1e-2 * np.abs(exp).max()          # one bf16 ulp near ties
    assert np.mean(got == exp) > 0.99
    mask = torch.from_numpy(mo.bf16(rng.normal(0, 1, (M, N)).astype(np.float32))).to(torch.bfloat16).cuda()
    _lib.call("fsx_mlp_gemm", ta, tb, 2, P(A), lda, P(Bm), ldb, M, N, K, None, P(c16), N, None, P(mask), st)
    exp = mo.bf16(np.where(mask.float().cpu().numpy() > 0, ref, 0).astype(np.float32))
    got = c16.float().cpu().numpy()
    assert np.mean(got == exp) > 0.99 and np.abs(got - exp).max() <= 1e-2 * np.abs(exp).max()


def _sync_params(net):
    return {k: v.cpu().numpy().astype(np.float64) for k, v in net.p.items()}


@pytest.mark.parametrize("ta,splits", [(1, 16), (1, 5), (0, 3)])
def test_gemm_splitk(gpu, ta, splits):
    """Split-K product = float64 product of the bf16 operands, deterministic (two runs equal),
    with chunks that do not divide K evenly."""
    import torch
    from fieldstitch_amd import _lib
    rng = np.random.default_rng(40 + splits)
    M, N, K = 288, 256, 8192 - 96
    A, a = _bf(torch, rng, (K, M) if ta else (M, K))
    Bm, b = _bf(torch, rng, (K, N))
    ref = (a.T if ta else a) @ b
    P = lambda t: C.c_void_p(t.data_ptr())   # noqa: E731
    st = C.c_void_p(_lib.stream_handle())
    sz = C.c_size_t()
    _lib.call("fsx_mlp_splitk_scratch", M, N, splits, C.byref(sz))
    scratch = torch.empty(sz.value // 4, dtype=torch.float32, device="cuda")
    outs = []
    for _ in range(2):
        c32 = torch.full((M, N + 3), -7.0, dtype=torch.float32, device="cuda")
        _lib.call("fsx_mlp_gemm_splitk", ta, 0, P(A), M if ta else K, P(Bm), N, M, N, K, P(c32), N + 3, P(scratch),
                  splits, st)
        outs.append(c32.cpu().numpy())
    np.testing.assert_array_equal(outs[0], outs[1])
    assert (outs[0][:, N:] == -7.0).all()                       # ldc padding untouched
    np.testing.assert_allclose(outs[0][:, :N], ref, rtol=1e-4, atol=1e-3)


def test_gradients_vs_oracle(gpu):
    import torch
    from fieldstitch_amd.filter_mlp import MatchFilterMLP, N_IN
    rng = np.random.default_rng(3)
    net = MatchFilterMLP(seed=5)
    nb = 512
    X = np.zeros((nb, 288), np.float32)
    X[:, :N_IN] = rng.normal(0, 1, (nb, N_IN))
    X = mo.bf16(X)
    y = (rng.random(nb) < 0.3).astype(np.float32)
    Xd = torch.from_numpy(X).to(torch.bfloat16).cuda()
    yd = torch.from_numpy(y).cuda()
    loss, g = net.gradients(Xd, yd)
    p = _sync_params(net)
    rloss, rg = mo.loss_and_grads(X.astype(np.float64), y, p, rounding=True)
    assert abs(float(loss) - rloss) <= 1e-4 * max(1.0, abs(rloss))
    for k in ("W1", "b1", "W2", "b2", "w3", "b3"):
        a = g[k].cpu().numpy().astype(np.float64).reshape(rg[k].shape)
        err = np.abs(a - rg[k]).max() / max(np.abs(rg[k]).max(), 1e-12)
        assert err < 2e-2, (k, err)
        # relative Frobenius error well below bf16 resolution on average
        assert np.linalg.norm(a - rg[k]) <= 1e-2 * np.linalg.norm(rg[k]) + 1e-12, k


def test_adam_step_vs_oracle(gpu):
    import torch
    from fieldstitch_amd.filter_mlp import MatchFilterMLP, N_IN
    rng = np.random.default_rng(4)
    net = MatchFilterMLP(seed=6, lr=1e-2)
    nb = 256
    X = np.zeros((nb, 288), np.float32)
    X[:, :N_IN] = rng.normal(0, 1, (nb, N_IN))
    Xd = torch.from_numpy(mo.bf16(X)).to(torch.bfloat16).cuda()
    yd = torch.from_numpy((rng.random(nb) < 0.5).astype(np.float32)).cuda()
    p0 = {k: v.cpu().numpy().copy() for k, v in net.p.items()}
    _, g = net.gradients(Xd, yd)
    g = {k: v.cpu().numpy().reshape(p0[k].shape).copy() for k, v in g.items()}
    net.step(Xd, yd)
    for k in p0:
        w, _, _ = mo.adam(p0[k], np.zeros_like(p0[k]), np.zeros_like(p0[k]), g[k], 1, lr=1e-2)
        np.testing.assert_allclose(net.p[k].cpu().numpy(), w, rtol=1e-5, atol=1e-6)
    assert torch.equal(net.W1_16, net.p["W1"].to(torch.bfloat16))


def test_training_learns_a_rule(gpu):
    """A rule the network can express (sign of a fixed projection of the inputs): the loss
    falls and held-out accuracy is high."""
    import torch
    from fieldstitch_amd.filter_mlp import MatchFilterMLP, N_IN
    rng = np.random.default_rng(8)
    n = 32768
    X = np.zeros((n, 288), np.float32)
    X[:, :N_IN] = rng.normal(0, 1, (n, N_IN))
    wdir = rng.normal(0, 1, N_IN)
    y = ((X[:, :N_IN] @ wdir) > 0.3).astype(np.float32)
    Xd = torch.from_numpy(mo.bf16(X)).to(torch.bfloat16).cuda()
    yd = torch.from_numpy(y).cuda()
    net = MatchFilterMLP(seed=1, lr=1e-3)
    losses = []
    for ep in range(6):
        for s in range(0, 28672, 4096):
            losses.append(float(net.step(Xd[s:s + 4096], yd[s:s + 4096])))
    p = net.predict(Xd[28672:]).cpu().numpy()
    acc = np.mean((p > 0.5) == (y[28672:] > 0.5))
    assert losses[-1] < 0.5 * losses[0] and acc > 0.9, (losses[0], losses[-1], acc)


@pytest.mark.parametrize("n", [1, 255, 1000, 70000])
def test_fused_inference(gpu, n):
    """predict's single fused launch against the training forward (GEMM epilogues + head) and
    the float64 oracle with the same bf16 roundings, on ragged row counts; the packed weights
    follow an Adam step."""
    import torch
    from fieldstitch_amd.filter_mlp import MatchFilterMLP, N_IN
    rng = np.random.default_rng(50 + n)
    net = MatchFilterMLP(seed=7, lr=1e-2)
    Xb = np.zeros((512, 288), np.float32)
    Xb[:, :N_IN] = rng.normal(0, 1, (512, N_IN))
    Xbd = torch.from_numpy(mo.bf16(Xb)).to(torch.bfloat16).cuda()
    net.step(Xbd, torch.from_numpy((rng.random(512) < 0.4).astype(np.float32)).cuda())
    for rep in range(2):
        X = np.zeros((n, 288), np.float32)
        X[:, :N_IN] = rng.normal(0, 1.5, (n, N_IN))
        X = mo.bf16(X)
        Xd = torch.from_numpy(X).to(torch.bfloat16).cuda()
        print("fused inference n=%d rep=%d" % (n, rep), flush=True)
        got = net.predict(Xd).cpu().numpy().astype(np.float64)
        fwd = net.forward(Xd).cpu().numpy().astype(np.float64)
        ref, _ = mo.forward(X.astype(np.float64), _sync_params(net), rounding=True)
        d_fwd = np.abs(got - fwd)
        d_ref = np.abs(got - ref)
        print("max |fused - forward| %.3g, exact %.4f; max |fused - oracle| %.3g" %
              (d_fwd.max(), np.mean(d_fwd == 0), d_ref.max()), flush=True)
        # a hidden unit may round to the other bf16 neighbour when the two float32 sums differ
        # in their last bits; that moves p by far less than 1e-2
        assert d_fwd.max() <= 1e-2 and d_ref.max() <= 1e-2
        assert np.mean(d_fwd <= 1e-5) > 0.95 and np.mean(d_ref <= 1e-4) > 0.95
        net.step(Xbd, torch.from_numpy((rng.random(512) < 0.4).astype(np.float32)).cuda())   # weights move
    assert net.predict(Xd[:0]).shape == (0,)
