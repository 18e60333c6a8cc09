"""The track oracle (recon.py:14-56 restated with scipy's connected_components) against a
pure-Python breadth-first reference, hand-built cases and flight-size invariants (CPU)."""
import numpy as np
import pytest

from oracle import tracks_oracle as O


def _random_case(seed, n_img=6, max_feat=12, n_pairs=10, n_match=8):
    rng = np.random.default_rng(seed)
    nfeat = rng.integers(0, max_feat + 1, n_img)
    nfeat[rng.integers(0, n_img)] = max_feat            # at least one photo with features
    pairs = []
    for _ in range(n_pairs):
        i, j = (int(x) for x in rng.integers(0, n_img, 2))
        if nfeat[i] == 0 or nfeat[j] == 0:
            continue
        m = int(rng.integers(0, n_match + 1))
        pairs.append((i, j, rng.integers(0, nfeat[i], m), rng.integers(0, nfeat[j], m)))
    return pairs, nfeat


def _eq(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3])) and a[3] == b[3]


@pytest.mark.parametrize("seed", range(40))
def test_oracle_equals_breadth_first_reference(seed):
    pairs, nfeat = _random_case(seed)
    assert _eq(O.build_tracks(pairs, nfeat, len(nfeat)), O.brute_force(pairs, nfeat, len(nfeat)))


def test_oracle_equals_reference_on_small_flight():
    pairs, nfeat = O.synthetic_flight_matches(n_img=16, nfeat=30, grid=4, seed=3)
    a = O.build_tracks(pairs, nfeat, 16)
    assert a[3] > 50 and _eq(a, O.brute_force(pairs, nfeat, 16))


def test_scipy_labels_follow_smallest_node():
    """build_tracks numbers tracks by scipy's component label; the GPU numbers them by their
    smallest node.  The two orders agree when scipy's labels rise with the smallest node."""
    import scipy.sparse as sp
    from scipy.sparse.csgraph import connected_components
    rng = np.random.default_rng(0)
    for n, e in ((50, 30), (2000, 1500), (100000, 90000)):
        u, v = rng.integers(0, n, e), rng.integers(0, n, e)
        _, lab = connected_components(sp.csr_matrix((np.ones(e), (u, v)), shape=(n, n)), directed=False)
        first = np.full(lab.max() + 1, n)
        np.minimum.at(first, lab, np.arange(n))
        assert np.all(np.diff(first) > 0)


def test_hand_built_cases():
    nf = [3, 3, 3, 2]
    # chain across photos 0-1-2: one track of three; a lone pair 0:2-3:1: a track of two
    r = O.build_tracks([(0, 1, [0], [1]), (1, 2, [1], [2]), (0, 3, [2], [1])], nf, 4)
    assert r[3] == 2
    assert r[0].tolist() == [0, 0, 0, 1, 1]
    assert r[1].tolist() == [0, 1, 2, 0, 3]
    assert r[2].tolist() == [0, 1, 2, 2, 1]
    # two features of photo 0 in one component: the whole component goes
    r = O.build_tracks([(0, 1, [0, 1], [0, 0]), (2, 3, [0], [0])], nf, 4)
    assert r[3] == 1 and r[1].tolist() == [2, 3]
    # a match of a feature with itself is a component of one: no track
    r = O.build_tracks([(1, 1, [2], [2])], nf, 4)
    assert r[3] == 0 and len(r[0]) == 0
    # tracks numbered by their smallest node, not by pair order
    r = O.build_tracks([(2, 3, [0], [0]), (0, 1, [2], [2])], nf, 4)
    assert r[1].tolist() == [0, 1, 2, 3]
    # no pairs, and pairs without matches
    assert O.build_tracks([], nf, 4)[3] == 0
    r = O.build_tracks([(0, 1, np.zeros(0, int), np.zeros(0, int))], nf, 4)
    assert r[3] == 0 and r[0].dtype == np.int64


def test_flight_invariants():
    pairs, nfeat = O.synthetic_flight_matches(n_img=64, nfeat=400, grid=8, seed=1)
    tid, img, feat, n = O.build_tracks(pairs, nfeat, 64)
    assert n > 1000 and tid[-1] == n - 1 and np.all(np.diff(tid) >= 0)
    lens = np.bincount(tid)
    assert lens.min() >= 2
    key = tid * 64 + img
    assert np.all(np.diff(key) > 0)                        # one photo per track, ascending
    node = np.array([0, *np.cumsum(nfeat)])[img] + feat
    first = np.searchsorted(tid, np.arange(n))
    assert np.all(np.diff(node[first]) > 0)                 # numbered by smallest node
