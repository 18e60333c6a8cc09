"""GPU matcher parity: plain FP32 and MFMA forms against the FP64 oracle (bit-exact match lists),
on synthetic RootSIFT-like uint8 descriptors with planted matches, guided (kc = 0 and kc = 16)
and brute force, with the edge cases of matching.py (queries outside B, single candidates,
fewer than two keypoints, empty sets)."""
import numpy as np
import pytest

from oracle import match_oracle as mo

pytestmark = pytest.mark.gpu


def test_mfma_layout(gpu):
    from fieldstitch_amd import matching
    out = matching.mfma_layout_probe()
    A = np.zeros((32, 16))
    Bm = np.zeros((16, 32))
    for lane in range(64):
        for e in range(8):
            k = (lane >> 5) * 8 + e
            A[lane & 31, k] = (lane & 31) + (k & 3)
            Bm[k, lane & 31] = ((lane & 31) * 3 + k) & 7
    Cm = A @ Bm
    for lane in range(64):
        for i in range(4):
            for j in range(4):
                assert out[lane, i * 4 + j] == Cm[i * 8 + (lane >> 5) * 4 + j, lane & 31]


def test_mfma_i8_layout(gpu):
    from fieldstitch_amd import matching
    out = matching.mfma_i8_layout_probe()
    # C[i][j] = sum over h, e of A(i, h, e) * B(j, h, e): both operands use the same k positions
    f = np.array([[[(i * 7 + h * 3 + e * 5) % 11 - 5 for e in range(16)] for h in range(2)] for i in range(32)])
    g = np.array([[[(j * 3 + h * 5 + e * 7) % 13 - 6 for e in range(16)] for h in range(2)] for j in range(32)])
    Cm = np.einsum("ihe,jhe->ij", f, g)
    for lane in range(64):
        for i in range(4):
            for j in range(4):
                assert out[lane, i * 4 + j] == Cm[i * 8 + (lane >> 5) * 4 + j, lane & 31]


def rootsift_like(rng, n):
    d = rng.gamma(0.6, 1.0, (n, 128))
    d /= d.sum(1, keepdims=True)
    return np.clip(np.sqrt(d) * 512, 0, 255).astype(np.uint8)


def make_pair(rng, na, nb, W=800, H=600, noise=10, off=(40.0, -25.0)):
    xy_b = np.c_[rng.uniform(0, W, nb), rng.uniform(0, H, nb)]
    db = rootsift_like(rng, nb)
    # A: half are copies of B keypoints (moved by -off), the rest random
    nt = min(na, nb) // 2
    src = rng.permutation(nb)[:nt]
    xy_a = np.r_[xy_b[src] - off + rng.normal(0, 1.0, (nt, 2)),
                 np.c_[rng.uniform(-100, W + 100, na - nt), rng.uniform(-100, H + 100, na - nt)]]
    da = np.r_[np.clip(db[src].astype(int) + rng.integers(-noise, noise + 1, (nt, 128)), 0, 255),
               rootsift_like(rng, na - nt)].astype(np.uint8)
    Hm = np.array([[1.0, 0, off[0]], [0, 1.0, off[1]], [0, 0, 1.0]])
    return da, xy_a, db, xy_b, Hm, (W, H)


def oracle_guided(da, xy_a, db, xy_b, Hm, wh, radius, kc, ratio):
    pred = mo.predict(xy_a, Hm)
    qs, cands = mo.guided_candidates(pred, xy_b, wh, radius, kc)
    q, t, _ = mo.match_fp64(da, db, ratio, qs, cands)
    return q, t


@pytest.mark.parametrize("form", ["plain", "mfma"])
@pytest.mark.parametrize("kc", [0, 16, 3])
def test_guided_vs_fp64_oracle(gpu, form, kc):
    from fieldstitch_amd.matching import PairMatcher
    rng = np.random.default_rng(7 + kc)
    images, pairs, Hms, expect = [], [], [], []
    for p in range(3):
        da, xy_a, db, xy_b, Hm, wh = make_pair(rng, 900 + 300 * p, 1100, W=700 + 50 * p)
        images += [(da, xy_a, wh), (db, xy_b, wh)]
        pairs.append((2 * p, 2 * p + 1))
        Hms.append(Hm)
        expect.append(oracle_guided(da, xy_a, db, xy_b, Hm, wh, 45.0, kc, 0.8))
    print("case guided", form, kc, flush=True)
    m = PairMatcher(images, pairs, Hms, radius=45.0, kc=kc, ratio=0.8)
    m.run(form)
    got = m.results()
    for (q, t), (qe, te) in zip(got, expect):
        assert len(qe) > 50
        np.testing.assert_array_equal(q, qe)
        np.testing.assert_array_equal(t, te)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_bruteforce_vs_fp64_oracle(gpu, form):
    from fieldstitch_amd import matching
    rng = np.random.default_rng(3)
    da, _, db, _, _, _ = make_pair(rng, 1500, 2100)
    print("case bf", form, flush=True)
    q, t = matching._ratio_match_bf(da.astype(np.float32), db.astype(np.float32), 0.8, form=form)
    qe, te, _ = mo.match_fp64(da, db, 0.8)
    np.testing.assert_array_equal(q, qe)
    np.testing.assert_array_equal(t, te)
    q2, t2 = mo.ratio_match_bf_f32(da.astype(np.float32), db.astype(np.float32), 0.8)
    np.testing.assert_array_equal(q, q2)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_edge_cases(gpu, form):
    from fieldstitch_amd import matching
    from fieldstitch_amd.matching import PairMatcher
    rng = np.random.default_rng(11)
    # fewer than two keypoints -> empty (matching.py:88-89)
    a = rootsift_like(rng, 1)
    b = rootsift_like(rng, 10)
    assert len(matching._ratio_match_bf(a, b, 0.8, form=form)[0]) == 0
    # a single candidate always passes; queries predicted outside B never match
    db = rootsift_like(rng, 2)
    xy_b = np.array([[100.0, 100.0], [400.0, 100.0]])
    da = db.copy()
    xy_a = np.array([[100.0, 100.0], [400.0, 100.0]])
    A = {"desc": da, "und": np.r_[xy_a, [[-500.0, 0.0]]], "wh": (500, 200)}
    A["desc"] = np.r_[da, rootsift_like(rng, 1)]
    B = {"desc": db, "und": xy_b, "wh": (500, 200)}
    q, t = matching._guided(A, B, np.eye(3), 30.0, 16, 0.8, form=form)
    assert list(q) == [0, 1] and list(t) == [0, 1]
    # empty keypoint set on either side, and a pair list mixing dense and guided pairs
    e = np.zeros((0, 128), np.uint8)
    ez = np.zeros((0, 2))
    m = PairMatcher([(e, ez, (10, 10)), (db, xy_b, (500, 200)), (da, xy_a, (500, 200))],
                    [(0, 1), (1, 0), (2, 1), (1, 2)], [np.eye(3), np.eye(3), np.eye(3), None], radius=30.0)
    m.run(form)
    r = m.results()
    assert len(r[0][0]) == 0 and len(r[1][0]) == 0
    assert list(r[2][0]) == [0, 1]
    qe, te, _ = mo.match_fp64(db, da, 0.8)
    np.testing.assert_array_equal(r[3][0], qe)


def test_forms_agree_on_flight(gpu):
    """Both forms on a realistic dense keypoint layout (many queries per cell)."""
    from fieldstitch_amd.matching import PairMatcher
    rng = np.random.default_rng(5)
    da, xy_a, db, xy_b, Hm, wh = make_pair(rng, 6000, 6000, W=2640, H=1978)
    m = PairMatcher([(da, xy_a, wh), (db, xy_b, wh)], [(0, 1), (1, 0)], [Hm, np.linalg.inv(Hm)], radius=60.0)
    m.run("plain")
    r0 = m.results()
    m.run("mfma")
    r1 = m.results()
    for (q0, t0), (q1, t1) in zip(r0, r1):
        np.testing.assert_array_equal(q0, q1)
        np.testing.assert_array_equal(t0, t1)
    qe, te = oracle_guided(da, xy_a, db, xy_b, Hm, wh, 60.0, 0, 0.8)
    np.testing.assert_array_equal(r1[0][0], qe)
    np.testing.assert_array_equal(r1[0][1], te)


@pytest.mark.parametrize("form", ["plain", "mfma"])
def test_graph_replays_equal_eager(gpu, form):
    """The launch sequence captured in a HIP graph gives the eager result on every replay, with
    eager work and synchronisations in between."""
    import torch
    from fieldstitch_amd.matching import PairMatcher
    rng = np.random.default_rng(31)
    imgs, pairs, Hms = [], [], []
    for i in range(3):
        da, xy_a, db, xy_b, Hm, wh = make_pair(rng, 1500, 1700)
        imgs += [(da, xy_a, wh), (db, xy_b, wh)]
        pairs.append((2 * i, 2 * i + 1))
        Hms.append(Hm)
    m = PairMatcher(imgs, pairs, Hms, radius=40.0, kc=0, ratio=0.8)
    m.run(form)
    ref = m.results()
    for rep in range(4):
        m.run(form, graph=True)
        torch.cuda.synchronize()
        got = m.results()
        for (q0, t0), (q1, t1) in zip(ref, got):
            assert np.array_equal(q0, q1) and np.array_equal(t0, t1), "replay %d differs" % rep
        m.run("plain" if form == "mfma" else "mfma")      # eager work in between
        torch.cuda.synchronize()


def test_persistent_workgroups_equal_oracle(gpu):
    """The MFMA contraction run by a fixed number of persistent workgroups (striding over all
    pairs' chunks) gives the one-workgroup-per-chunk result and the FP64 oracle's, on pairs of
    uneven sizes including one whose prior puts every query outside the target photo (all its
    chunks inactive); then the autotuner's choice is one of the candidates."""
    from fieldstitch_amd import autotune
    from fieldstitch_amd.matching import PairMatcher
    rng = np.random.default_rng(41)
    images, pairs, Hms, expect = [], [], [], []
    for p, (na, nb) in enumerate(((300, 900), (2500, 1200), (1100, 3000), (700, 700))):
        da, xy_a, db, xy_b, Hm, wh = make_pair(rng, na, nb, W=700 + 40 * p)
        if p == 3:
            Hm = Hm.copy()
            Hm[0, 2] += 5000.0           # every predicted position far outside photo b
        images += [(da, xy_a, wh), (db, xy_b, wh)]
        pairs.append((2 * p, 2 * p + 1))
        Hms.append(Hm)
        expect.append(oracle_guided(da, xy_a, db, xy_b, Hm, wh, 45.0, 0, 0.8))
    pairs.append((2, 5))                 # photos shared between pairs
    Hms.append(np.eye(3))
    expect.append(oracle_guided(images[2][0], images[2][1], images[5][0], images[5][1], np.eye(3), images[5][2],
                                45.0, 0, 0.8))
    m = PairMatcher(images, pairs, Hms, radius=45.0, kc=0, ratio=0.8)
    old = autotune.get(autotune.MATCH_BLOCKS)
    got = {}
    try:
        for wg in (0, 256, 4096):
            autotune.set(autotune.MATCH_BLOCKS, wg)
            print("case persistent workgroups", wg, flush=True)
            m.run("mfma")
            got[wg] = m.results()
    finally:
        autotune.set(autotune.MATCH_BLOCKS, old)
    assert len(expect[3][0]) == 0 and len(expect[1][0]) > 50
    for wg, res in got.items():
        for (q, t), (qe, te) in zip(res, expect):
            np.testing.assert_array_equal(q, qe)
            np.testing.assert_array_equal(t, te)
    try:
        tuned = autotune.tune_matcher(m, "mfma", reps=2)
        assert tuned["chosen"] in autotune.CANDIDATES[autotune.MATCH_BLOCKS]
        m.run("mfma")
        for (q, t), (qe, te) in zip(m.results(), expect):
            np.testing.assert_array_equal(q, qe)
            np.testing.assert_array_equal(t, te)
    finally:
        autotune.set(autotune.MATCH_BLOCKS, old)
