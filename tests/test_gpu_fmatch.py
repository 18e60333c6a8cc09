"""Float-descriptor brute-force matcher (BF16 / FP8 screening + float64 certificate) against
the FP64 oracle: the match lists must be identical."""
import numpy as np
import pytest

from fieldstitch_amd import synth
from oracle import match_oracle as mo

pytestmark = pytest.mark.gpu


def test_fp8_layout(gpu):
    from fieldstitch_amd import matching
    out = matching.fp8_layout_probe()
    f = np.array([[[(i * 7 + h * 3 + e * 5) % 11 - 5 for e in range(32)] for h in range(2)] for i in range(32)])
    g = np.array([[[(j * 3 + h * 5 + e * 7) % 13 - 6 for e in range(32)] for h in range(2)] for j in range(32)])
    Cm = np.einsum("ihe,jhe->ij", f, g).astype(np.float64)
    for lane in range(64):
        for i in range(4):
            for j in range(4):
                assert out[lane, i * 4 + j] == Cm[i * 8 + (lane >> 5) * 4 + j, lane & 31]


def test_fp8_scale_layout(gpu):
    """E8M0 scales: lane l's scale applies to A row l % 32 (its K block) and to B column l % 32."""
    from fieldstitch_amd import matching
    out = matching.fp8_scale_probe()
    f = np.array([[[(i * 7 + h * 3 + e * 5) % 11 - 5 for e in range(32)] for h in range(2)] for i in range(32)])
    g = np.array([[[(j * 3 + h * 5 + e * 7) % 13 - 6 for e in range(32)] for h in range(2)] for j in range(32)])
    Cm = np.einsum("ihe,jhe->ij", f, g).astype(np.float64)
    sa = 2.0 ** (np.arange(32) % 5 - 2)
    sb = 2.0 ** (np.arange(32) % 3 - 1)
    Cm = Cm * sa[:, None] * sb[None, :]
    for lane in range(64):
        for i in range(4):
            for j in range(4):
                assert out[lane, i * 4 + j] == Cm[i * 8 + (lane >> 5) * 4 + j, lane & 31]


def test_fp8_hardware_conversion_is_rne_e4m3(gpu):
    """The FP8 quantization's hardware conversion gives the software round-to-nearest-even
    e4m3 code for every float in [-448, 448] (all 2^31 + 2 bit patterns)."""
    from fieldstitch_amd import matching
    bad, first = matching.fp8_cvt_probe()
    assert bad == 0, (bad, None if first is None else hex(first))


def _oracle(da, db, ratio):
    q, t, _ = mo.match_fp64(da, db, ratio)
    return q, t


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
@pytest.mark.parametrize("K", [8, 2, 0])
def test_float_bruteforce_vs_fp64(gpu, precision, K):
    from fieldstitch_amd.matching import FloatMatcher
    rng = np.random.default_rng(100 + K)
    sizes = [(1500, 1800), (2, 700), (900, 2), (1, 50), (640, 1200)]
    sets, pairs, ref = [], [], []
    for na, nb in sizes:
        da, db, _ = synth.descriptor_pair(rng, na, nb)
        pairs.append((len(sets), len(sets) + 1))
        sets += [da, db]
        ref.append(_oracle(da, db, 0.8))
    print("precision", precision, "K", K, flush=True)
    fm = FloatMatcher(sets, pairs, ratio=0.8, precision=precision, K=K)
    fm.run()
    got = fm.results()
    for (q, t), (rq, rt), (na, nb) in zip(got, ref, sizes):
        assert np.array_equal(q, rq) and np.array_equal(t, rt), (na, nb)
    st = fm.stats()
    if K == 0:       # exact float64 pass for every query of the pairs with >= 2 on each side
        assert st["fallback_queries"] == sum(na for na, nb in sizes if na >= 2 and nb >= 2)
    else:
        assert st["rescored_candidates"] > 0 and np.isfinite(st["max_abs_distance_error"])


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
def test_float_ties_and_duplicates(gpu, precision):
    """Exact duplicate targets (equal distances: best by index, b2 == b1 rejects) and exact
    copies of queries (distance 0)."""
    from fieldstitch_amd.matching import FloatMatcher
    rng = np.random.default_rng(7)
    da, db, _ = synth.descriptor_pair(rng, 800, 900)
    db[100] = db[200]                    # two equal targets
    db[300] = da[5]                      # exact copy of query 5
    db[301] = da[6]
    db[302] = da[6]                      # query 6: two exact copies -> b1 == b2 -> rejected
    da[7] = da[8]                        # two equal queries compete for one target
    fm = FloatMatcher([da, db], [(0, 1)], ratio=0.8, precision=precision)
    fm.run()
    (q, t), = fm.results()
    rq, rt = _oracle(da, db, 0.8)
    assert np.array_equal(q, rq) and np.array_equal(t, rt)


def test_ratio_match_bf_dispatch(gpu):
    """_ratio_match_bf: uint8-valued inputs take the integer path, float inputs the float path;
    both equal the FP64 oracle."""
    from fieldstitch_amd.matching import _ratio_match_bf
    rng = np.random.default_rng(9)
    da, db, _ = synth.descriptor_pair(rng, 700, 800)
    q, t = _ratio_match_bf(da, db, 0.8)
    rq, rt = _oracle(da, db, 0.8)
    assert np.array_equal(q, rq) and np.array_equal(t, rt)
    ua, ub = np.clip(np.rint(da / 2), 0, 255).astype(np.float32), np.clip(np.rint(db / 2), 0, 255).astype(np.float32)
    q, t = _ratio_match_bf(ua, ub, 0.8)
    rq, rt = _oracle(ua, ub, 0.8)
    assert np.array_equal(q, rq) and np.array_equal(t, rt)


def test_non_finite_descriptors_rejected(gpu):
    from fieldstitch_amd.matching import FloatMatcher
    rng = np.random.default_rng(1)
    a = rng.random((50, 128)).astype(np.float32)
    b = rng.random((60, 128)).astype(np.float32)
    b[7, 3] = np.nan
    with pytest.raises(ValueError):
        FloatMatcher([a, b], [(0, 1)])
