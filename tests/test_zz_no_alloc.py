"""fsx_match_batch_noalloc (the matcher for callers that cannot allocate) through the C ABI
against fsx_match_batch on the same inputs: bit-identical counts and (q, t) lists, guided
(kc = 0 and kc = 16) and brute-force pairs, both contraction forms, on two calls in a row, the
second smaller than the first (the shared scratch keeps the first call's contents)."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def rootsift_like(rng, n):
    d = rng.gamma(0.6, 1.0, (n, 128))
    d /= d.sum(1, keepdims=True)
    return np.clip(np.sqrt(d) * 512, 0, 255).astype(np.uint8)


def make_images(rng, sizes, W=900, H=700):
    """Keypoint sets that overlap: half of image i's points are a prefix of one shared point
    list shifted by (30 i, -20 i) with jitter (descriptors perturbed copies), the rest fresh."""
    base_xy = np.c_[rng.uniform(0, W, 4000), rng.uniform(0, H, 4000)]
    base_d = rootsift_like(rng, 4000)
    order = rng.permutation(4000)
    out = []
    for i, n in enumerate(sizes):
        k = n // 2
        src = order[:k]          # every photo sees a prefix of the same points
        xy = np.r_[base_xy[src] + (30.0 * i, -20.0 * i) + rng.normal(0, 1.0, (k, 2)),
                   np.c_[rng.uniform(0, W, n - k), rng.uniform(0, H, n - k)]]
        d = np.r_[np.clip(base_d[src].astype(int) + rng.integers(-8, 9, (k, 128)), 0, 255),
                  rootsift_like(rng, n - k)].astype(np.uint8)
        out.append((d, xy, (W, H)))
    return out


def shift(i, j):
    return np.array([[1.0, 0, 30.0 * (j - i)], [0, 1.0, -20.0 * (j - i)], [0, 0, 1.0]])


def noalloc(images, pairs, Hm, radius, kc, ratio, form):
    import torch
    from fieldstitch_amd import _lib
    from fieldstitch_amd.matching import FORMS
    n_img, n_pair = len(images), len(pairs)
    desc = torch.from_numpy(np.concatenate([d for d, _, _ in images])).cuda()
    xy = torch.from_numpy(np.ascontiguousarray(np.concatenate([x for _, x, _ in images]), np.float64)).cuda()
    nk = [len(d) for d, _, _ in images]
    cap = sum(min(nk[a], nk[b]) for a, b in pairs)
    q = torch.full((max(cap, 1),), -7, dtype=torch.int32, device="cuda")
    t = torch.full((max(cap, 1),), -7, dtype=torch.int32, device="cuda")
    counts = torch.full((n_pair,), -7, dtype=torch.int32, device="cuda")
    flags = torch.full((4,), -7, dtype=torch.int32, device="cuda")
    dense = [1 if h is None else 0 for h in Hm]
    hm = np.ascontiguousarray(np.stack([np.eye(3) if h is None else np.asarray(h, np.float64) for h in Hm]).reshape(-1))
    offs = (C.c_longlong * n_pair)()
    ci = lambda v: (C.c_int * len(v))(*v)   # noqa: E731
    P = C.c_void_p
    _lib.call("fsx_match_batch_noalloc", n_img, ci(nk), ci([w for _, _, (w, _) in images]),
              ci([h for _, _, (_, h) in images]), P(desc.data_ptr()), P(xy.data_ptr()), n_pair,
              ci([a for a, _ in pairs]), ci([b for _, b in pairs]), ci(dense),
              hm.ctypes.data_as(C.POINTER(C.c_double)), float(radius), int(kc), float(ratio), FORMS[form],
              P(q.data_ptr()), P(t.data_ptr()), P(counts.data_ptr()), P(flags.data_ptr()), offs,
              P(_lib.stream_handle()))
    qh, th, ch, fh = q.cpu().numpy(), t.cpu().numpy(), counts.cpu().numpy(), flags.cpu().numpy()
    assert fh[2] == 0
    return [(qh[offs[p]:offs[p] + ch[p]].astype(np.int64), th[offs[p]:offs[p] + ch[p]].astype(np.int64))
            for p in range(n_pair)]


def with_workspace(images, pairs, Hm, radius, kc, ratio, form):
    from fieldstitch_amd.matching import PairMatcher
    pm = PairMatcher(images, pairs, Hm, radius=radius, kc=kc, ratio=ratio)
    pm.run(form)
    return pm.results()


@pytest.mark.parametrize("form", ["mfma", "plain"])
@pytest.mark.parametrize("kc", [0, 16])
def test_noalloc_equals_workspace_entry_point(gpu, form, kc):
    rng = np.random.default_rng(1234 + kc)
    calls = [
        # first call: three photos, two guided pairs and one brute-force pair
        (make_images(rng, [3000, 2600, 1900]), [(0, 1), (1, 2), (0, 2)], "g,g,d"),
        # second call, smaller: two photos, one guided pair
        (make_images(rng, [900, 700]), [(1, 0)], "g"),
    ]
    for images, pairs, kinds in calls:
        Hm = [shift(a, b) if k == "g" else None for (a, b), k in zip(pairs, kinds.split(","))]
        got = noalloc(images, pairs, Hm, 40.0, kc, 0.8, form)
        want = with_workspace(images, pairs, Hm, 40.0, kc, 0.8, form)
        assert sum(len(w[0]) for w in want) > 100
        for g, w in zip(got, want):
            np.testing.assert_array_equal(g[0], w[0])
            np.testing.assert_array_equal(g[1], w[1])
