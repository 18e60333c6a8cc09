"""The chain's device-side list moves: fsx_gather_lists against a numpy restatement (selection,
keep masks, positions with out-of-range rows, the capacity cut, kept counts) and fsx_match_pack
against the per-query output slots it replaces."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _gather_ref(src_off, cnt, dst_off, cap, base_a, base_b, q, t, keep, xy):
    n_out = int(max([d + c for d, c in zip(dst_off, cnt)] + [0]))
    qd = np.full(max(n_out, cap), -7, np.int32)
    td = np.full(max(n_out, cap), -7, np.int32)
    pa = np.full((max(n_out, cap), 2), -7.0)
    pb = np.full((max(n_out, cap), 2), -7.0)
    kept = []
    for L in range(len(cnt)):
        rows = [r for r in range(src_off[L], src_off[L] + cnt[L]) if keep is None or keep[r]]
        kept.append(len(rows))
        for k, r in enumerate(rows):
            d = dst_off[L] + k
            if d >= cap:
                continue
            qd[d], td[d] = q[r], t[r]
            ra, rb = base_a[L] + q[r], base_b[L] + t[r]
            pa[d] = xy[ra] if 0 <= ra < len(xy) else np.nan
            pb[d] = xy[rb] if 0 <= rb < len(xy) else np.nan
    return qd[:cap], td[:cap], pa[:cap], pb[:cap], np.array(kept)


@pytest.mark.parametrize("with_keep", [False, True])
def test_gather_lists_equals_restatement(gpu, with_keep):
    import torch
    from fieldstitch_amd import _lib
    rng = np.random.default_rng(40 + with_keep)
    n_src = 5000
    q = rng.integers(-2, 900, n_src).astype(np.int32)           # a few rows out of range
    t = rng.integers(0, 1200, n_src).astype(np.int32)
    xy = rng.normal(size=(2000, 2)) * 100
    n_list = 37
    cnt = rng.integers(0, 400, n_list)
    cnt[3] = 0
    cnt[7] = 300                                                  # longer than one block round
    src_off = rng.integers(0, n_src - 400, n_list)
    base_a = rng.integers(0, 1900, n_list)            # base + q passes the 2000 rows for some
    base_b = rng.integers(0, 800, n_list)
    keep = (rng.random(n_src) < 0.6).astype(np.uint8) if with_keep else None
    kc = np.array([sum(1 for r in range(s, s + c) if keep is None or keep[r]) for s, c in zip(src_off, cnt)])
    dst_off = np.r_[0, np.cumsum(kc)[:-1]]
    cap = int(kc.sum()) - 25                                      # the last rows are cut
    want = _gather_ref(src_off, cnt, dst_off, cap, base_a, base_b, q, t, keep, xy)
    d = lambda a, dt: torch.from_numpy(np.ascontiguousarray(np.asarray(a, dt))).cuda()
    qd = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    td = torch.full((cap,), -7, dtype=torch.int32, device="cuda")
    pa = torch.full((cap, 2), -7.0, dtype=torch.float64, device="cuda")
    pb = torch.full((cap, 2), -7.0, dtype=torch.float64, device="cuda")
    kept = torch.zeros(n_list, dtype=torch.int32, device="cuda")
    args = [d(src_off, np.int64), d(cnt, np.int32), d(dst_off, np.int64), None, d(base_a, np.int64),
            d(base_b, np.int64), d(q, np.int32), d(t, np.int32), d(keep, np.uint8) if with_keep else None,
            d(xy, np.float64)]
    P = lambda x: None if x is None else C.c_void_p(x.data_ptr())
    print("case gather_lists keep", with_keep, flush=True)
    _lib.call("fsx_gather_lists", n_list, P(args[0]), P(args[1]), P(args[2]), cap, P(args[4]), P(args[5]),
              P(args[6]), P(args[7]), P(args[8]), P(args[9]), len(xy), P(qd), P(td), P(pa), P(pb), P(kept),
              C.c_void_p(_lib.stream_handle()))
    torch.cuda.synchronize()
    np.testing.assert_array_equal(qd.cpu().numpy(), want[0])
    np.testing.assert_array_equal(td.cpu().numpy(), want[1])
    np.testing.assert_array_equal(pa.cpu().numpy(), want[2])          # NaN == NaN for assert_array_equal
    np.testing.assert_array_equal(pb.cpu().numpy(), want[3])
    np.testing.assert_array_equal(kept.cpu().numpy(), want[4])
    assert np.isnan(want[2]).any()                                # out-of-range rows were exercised


def test_match_pack_equals_output_slots(gpu):
    import torch
    from fieldstitch_amd import _lib
    from fieldstitch_amd.matching import PairMatcher
    from test_gpu_match import make_pair
    rng = np.random.default_rng(44)
    images, pairs, Hms = [], [], []
    for p in range(4):
        da, xy_a, db, xy_b, Hm, wh = make_pair(rng, 700 + 200 * p, 900)
        images += [(da, xy_a, wh), (db, xy_b, wh)]
        pairs.append((2 * p, 2 * p + 1))
        Hms.append(Hm)
    m = PairMatcher(images, pairs, Hms, radius=45.0, kc=0, ratio=0.8)
    print("case match_pack", flush=True)
    m.run("mfma")
    counts = (C.c_int * m.n_pair)()
    q = np.zeros(max(m.capacity, 1), np.int32)
    t = np.zeros(max(m.capacity, 1), np.int32)
    flags = (C.c_int * 4)()
    _lib.call("fsx_match_results", m.n_img, m._nk, m._W, m._H, m.n_pair, m._pa, m._pb, m.radius,
              C.c_void_p(m.ws.data_ptr()), counts, q.ctypes.data_as(C.POINTER(C.c_int)),
              t.ctypes.data_as(C.POINTER(C.c_int)), flags, C.c_void_p(_lib.stream_handle()))
    cnt, off, qd, td = m.results_dev()
    qd, td = qd.cpu().numpy(), td.cpu().numpy()
    assert list(cnt) == list(counts) and off[-1] == len(qd) and min(cnt) > 50
    for p in range(m.n_pair):
        o = m._offsets[p]
        np.testing.assert_array_equal(qd[off[p]:off[p + 1]], q[o:o + counts[p]])
        np.testing.assert_array_equal(td[off[p]:off[p + 1]], t[o:o + counts[p]])
    for (qa, ta), p in zip(m.results(), range(m.n_pair)):
        np.testing.assert_array_equal(qa, qd[off[p]:off[p + 1]])
        np.testing.assert_array_equal(ta, td[off[p]:off[p + 1]])
