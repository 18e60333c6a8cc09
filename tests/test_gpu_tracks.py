"""GPU track building (fieldstitch_amd.tracks, recon.py:14-56) against the oracle: bit-exact
tid / img / feat arrays and track counts on random small cases, hand-built cases, a 400-photo
flight's match lists, empty inputs and index errors."""
import numpy as np
import pytest

from oracle import tracks_oracle as O

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3])) and a[3] == b[3]


def _random_case(seed, n_img, max_feat, n_pairs, n_match):
    rng = np.random.default_rng(seed)
    nfeat = rng.integers(0, max_feat + 1, n_img)
    nfeat[0] = max_feat
    pairs = []
    for _ in range(n_pairs):
        i, j = (int(x) for x in rng.integers(0, n_img, 2))
        if nfeat[i] == 0 or nfeat[j] == 0:
            continue
        m = int(rng.integers(0, n_match + 1))
        pairs.append((i, j, rng.integers(0, nfeat[i], m), rng.integers(0, nfeat[j], m)))
    return pairs, nfeat


@pytest.mark.parametrize("seed", range(12))
def test_random_cases_match_oracle(gpu, seed):
    from fieldstitch_amd import tracks
    shape = [(6, 12, 10, 8), (30, 200, 120, 150), (80, 3000, 400, 2000), (50, 500, 200, 30)][seed % 4]
    pairs, nfeat = _random_case(seed, *shape)
    print("tracks random seed=%d shape=%s" % (seed, shape), flush=True)
    g = tracks.build_tracks(pairs, nfeat, len(nfeat))
    assert _eq(g, O.build_tracks(pairs, nfeat, len(nfeat)))
    for x in g[:3]:
        assert x.dtype == np.int64


def test_hand_built_cases(gpu):
    from fieldstitch_amd import tracks
    nf = [3, 3, 3, 2]
    cases = [[(0, 1, [0], [1]), (1, 2, [1], [2]), (0, 3, [2], [1])],
             [(0, 1, [0, 1], [0, 0]), (2, 3, [0], [0])],
             [(1, 1, [2], [2])],
             [(2, 3, [0], [0]), (0, 1, [2], [2])],
             [(0, 1, np.zeros(0, int), np.zeros(0, int))]]
    for c in cases:
        c = [(i, j, np.asarray(q), np.asarray(t)) for i, j, q, t in c]
        assert _eq(tracks.build_tracks(c, nf, 4), O.build_tracks(c, nf, 4))
    r = tracks.build_tracks([], nf, 4)
    assert r[3] == 0 and len(r[0]) == 0


def test_flight_matches_equal_oracle_and_repeat(gpu):
    """400 photos x 8000 features, 1482 pairs, 3.6 M matches with 2 % wrong ones: equal to the
    oracle, and equal again on a second run (union order differs run to run)."""
    from fieldstitch_amd import tracks
    pairs, nfeat = O.synthetic_flight_matches()
    print("tracks flight: %d pairs" % len(pairs), flush=True)
    ref = O.build_tracks(pairs, nfeat, 400)
    a = tracks.build_tracks(pairs, nfeat, 400)
    assert _eq(a, ref)
    assert _eq(tracks.build_tracks(pairs, nfeat, 400), ref)


def test_index_errors_raise(gpu):
    from fieldstitch_amd import tracks
    with pytest.raises(ValueError):
        tracks.build_tracks([(0, 1, np.array([3]), np.array([0]))], [3, 3], 2)
    with pytest.raises(ValueError):
        tracks.build_tracks([(0, 2, np.array([0]), np.array([0]))], [3, 3], 2)
    with pytest.raises(ValueError):
        tracks.build_tracks([(0, 1, np.array([0]), np.array([0]))], [3, 3], 1)
    # the builder is usable after an error
    r = tracks.build_tracks([(0, 1, np.array([0]), np.array([2]))], [3, 3], 2)
    assert r[3] == 1 and r[2].tolist() == [0, 2]


def test_long_tracks_and_many_photos(gpu):
    """Long tracks (a 150-photo chain kept, a 120-photo chain with a repeated photo dropped),
    components with more members than photos, and more photos than the kernels stage in LDS
    (3000 photos)."""
    from fieldstitch_amd import tracks
    rng = np.random.default_rng(5)
    n_img = 200
    nfeat = np.full(n_img, 10, np.int64)
    order = rng.permutation(n_img)
    pairs = []
    for a, b in zip(order[:149], order[1:150]):          # one feature per photo, 150 photos
        pairs.append((int(a), int(b), np.array([3]), np.array([3])))
    chain = rng.permutation(n_img)[:120]
    for a, b in zip(chain[:-1], chain[1:]):
        pairs.append((int(a), int(b), np.array([7]), np.array([7])))
    pairs.append((int(chain[0]), int(chain[0]), np.array([7]), np.array([8])))   # photo repeats
    pairs.append((0, 1, np.array([0, 1]), np.array([0, 1])))
    print("tracks long chains", flush=True)
    g = tracks.build_tracks(pairs, nfeat, n_img)
    ref = O.build_tracks(pairs, nfeat, n_img)
    assert _eq(g, ref) and np.bincount(g[0]).max() == 150
    print("tracks 3000 photos", flush=True)
    pairs, nfeat = _random_case(9, 3000, 6, 6000, 4)
    assert _eq(tracks.build_tracks(pairs, nfeat, 3000), O.build_tracks(pairs, nfeat, 3000))
    # one dense cluster: more members than photos
    pairs = [(i % 4, (i + 1) % 4, np.arange(6), np.roll(np.arange(6), 1)) for i in range(4)]
    assert _eq(tracks.build_tracks(pairs, [6] * 4, 4), O.build_tracks(pairs, [6] * 4, 4))
