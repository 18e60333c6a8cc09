"""The C ABI: libfsx.so loads, exports every function include/fsx.h declares, and its host-side
helpers (octave plan, Gaussian taps) agree with the oracle.  No GPU needed."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    txt = open(os.path.join(ROOT, "include", "fsx.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\bint\s+(fsx_\w+)\s*\(", txt)))


def test_header_declares_functions():
    names = declared_functions()
    assert "fsx_sift_detect" in names and "fsx_last_error" in names


def test_library_exports_all_declared():
    from fieldstitch_amd import _lib
    L = _lib.load()
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the Python binding knows their signatures
    assert not [n for n in declared_functions() if n not in _lib.SIGNATURES]


def test_error_path_sets_message():
    from fieldstitch_amd import _lib
    with pytest.raises(_lib.FsxError, match="bad image size"):
        _lib.call("fsx_sift_plan", 0, 10, C.byref(C.c_int()), None, None)


@pytest.mark.parametrize("H,W", [(912, 1368), (1978, 2640), (989, 1320), (240, 320), (64, 80), (7, 9)])
def test_plan_matches_oracle(H, W):
    from fieldstitch_amd import gsift
    from oracle import sift_oracle as so
    assert gsift.plan(H, W) == so.plan(H, W)


def test_taps_match_oracle_and_numpy():
    from fieldstitch_amd import gsift
    from oracle import sift_oracle as so
    sig = so.sigmas()
    for i in range(6):
        k = gsift.taps(i)
        np.testing.assert_array_equal(k, so.gauss_kernel(sig[i]))
        # the reference's formula evaluated by numpy itself (gsift.py:52-56)
        n = int(round(sig[i] * 8 + 1)) | 1
        x = np.arange(n, dtype=np.float64) - (n - 1) / 2.0
        g = np.exp(-x * x / (2 * sig[i] * sig[i]))
        np.testing.assert_array_equal(k, (g / g.sum()).astype(np.float32))
    assert [len(gsift.taps(i)) for i in range(6)] == [11, 11, 13, 17, 21, 27]


def test_workspace_bytes_rejects_bad_caps():
    from fieldstitch_amd import _lib
    n = C.c_size_t()
    with pytest.raises(_lib.FsxError):
        _lib.call("fsx_sift_workspace_bytes", 100, 100, 1000, 1000, C.byref(n))
    _lib.call("fsx_sift_workspace_bytes", 100, 100, 1000, 1024, C.byref(n))
    assert n.value > 4 * 200 * 200 * 9


def _noalloc_args(nk, pairs, wh=(1000, 800)):
    n = len(nk)
    ci = lambda v: (C.c_int * max(len(v), 1))(*v)   # noqa: E731
    hm = (C.c_double * (9 * max(len(pairs), 1)))()
    fake = C.c_void_p(256)      # never dereferenced: every rejection below happens on the host
    return (n, ci(nk), ci([wh[0]] * n), ci([wh[1]] * n), fake, fake, len(pairs), ci([a for a, _ in pairs]),
            ci([b for _, b in pairs]), ci([1] * len(pairs)), hm, 40.0, 0, 0.8, 1, fake, fake, fake, fake,
            (C.c_longlong * max(len(pairs), 1))(), None)


def test_noalloc_matcher_rejects_what_exceeds_its_limits():
    """fsx_match_batch_noalloc refuses too many pairs / images and batches whose workspace
    exceeds its 32 MiB scratch, with a message, before any device work."""
    from fieldstitch_amd import _lib
    with pytest.raises(_lib.FsxError, match="n_pair must be in"):
        _lib.call("fsx_match_batch_noalloc", *_noalloc_args([100, 100], [(0, 1)] * 257))
    with pytest.raises(_lib.FsxError, match="n_img must be in"):
        _lib.call("fsx_match_batch_noalloc", *_noalloc_args([10] * 65, [(0, 1)]))
    with pytest.raises(_lib.FsxError, match="bytes of matcher scratch"):
        _lib.call("fsx_match_batch_noalloc", *_noalloc_args([600000, 600000], [(0, 1)]))
    with pytest.raises(_lib.FsxError, match="null pointer"):
        a = list(_noalloc_args([100, 100], [(0, 1)]))
        a[15] = None
        _lib.call("fsx_match_batch_noalloc", *a)


def test_tune_keys_validated():
    """Launch-shape tunables: accepted values round-trip, others are refused with a message
    (host code only)."""
    import ctypes as C
    from fieldstitch_amd import _lib, autotune
    for key, vals in autotune.CANDIDATES.items():
        old = autotune.get(key)
        try:
            for v in vals:
                autotune.set(key, v)
                assert autotune.get(key) == v
        finally:
            autotune.set(key, old)
    assert autotune.get(autotune.BLUR_ROWS) == 16
    L = _lib.load()
    for key, bad in ((autotune.BLUR_ROWS, 12), (autotune.FSCREEN_TILE, 96), (autotune.FSCREEN_QB, 3),
                     (autotune.MATCH_BLOCKS, 100), (autotune.MATCH_BLOCKS, -1), (6, 1)):
        assert L.fsx_tune_set(key, bad) != 0
    assert autotune.get(autotune.MATCH_BLOCKS) == 0      # default: one workgroup per chunk
    v = C.c_int()
    assert L.fsx_tune_get(6, C.byref(v)) != 0
