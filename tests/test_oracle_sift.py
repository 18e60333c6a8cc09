"""Pins of the SIFT oracle (oracle/sift_oracle.c) before it is trusted as the GPU checker:
the reference's own CPU outputs (golden vectors from app/planesite/gsift.py), library routines
(scipy.ndimage.correlate1d / maximum_filter), closed forms and invariants."""
import os

import numpy as np
import pytest
import scipy.ndimage as ndi

from oracle import sift_oracle as so
from fieldstitch_amd import synth

GOLD = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ["tiny_cv", "c1_crop_cv", "c1_crop_pipe", "c1_full_cv", "c1_full2_cv"]


def golden(name):
    return np.load(os.path.join(GOLD, "gsift_%s.npz" % name))


def golden_image(name, g):
    if "img" in g:
        return g["img"]
    img1, img2, _ = synth.field_pair(1)
    return img1 if name == "c1_full_cv" else img2


def test_golden_inputs_regenerate():
    import hashlib
    img1, img2, _ = synth.field_pair(1)
    for name, img in (("c1_full_cv", img1), ("c1_full2_cv", img2)):
        assert hashlib.sha256(img.tobytes()).hexdigest() == str(golden(name)["img_sha"])


# ---- S2: taps, sigmas (closed form) --------------------------------------------------
def test_sigmas_and_tap_counts():
    s = so.sigmas()
    np.testing.assert_allclose(s, [1.2490, 1.2263, 1.5450, 1.9466, 2.4525, 3.0900], atol=1e-4)
    assert [len(so.gauss_kernel(x)) for x in s] == [11, 11, 13, 17, 21, 27]
    for x in s:
        k = so.gauss_kernel(x).astype(np.float64)
        assert abs(k.sum() - 1.0) < 1e-6 and np.array_equal(k, k[::-1])


# ---- S4: blur == scipy.ndimage.correlate1d(mirror) twice, bit for bit ------------------
@pytest.mark.parametrize("shape", [(14, 15), (31, 50), (64, 33), (100, 14)])
def test_blur_bitexact_vs_scipy(shape):
    rng = np.random.default_rng(shape[0] * 7 + shape[1])
    img = rng.random(shape).astype(np.float32)
    for s in so.sigmas():
        k = so.gauss_kernel(s)
        ref = ndi.correlate1d(ndi.correlate1d(img, k, axis=0, mode="mirror"), k, axis=1, mode="mirror")
        np.testing.assert_array_equal(so.blur(img, k), ref)


# ---- S3: upsample closed form ------------------------------------------------------------
def test_up2_closed_form():
    g = np.array([[0, 255], [51, 102]], np.uint8)
    a = g.astype(np.float32) * np.float32(1.0 / 255.0)
    out = so.up2_u8(g)
    assert out.shape == (4, 4)
    # corner (0,0): rows clamp -> a00, cols clamp -> a00
    assert out[0, 0] == a[0, 0]
    # (1,1): 0.75 of (0.75 a00 + 0.25 a10) + 0.25 of (0.75 a01 + 0.25 a11)
    t0 = np.float32(0.75) * a[0, 0] + np.float32(0.25) * a[1, 0]
    t1 = np.float32(0.75) * a[0, 1] + np.float32(0.25) * a[1, 1]
    assert out[1, 1] == np.float32(0.75) * t0 + np.float32(0.25) * t1
    # constant image stays constant
    assert np.all(so.up2_u8(np.full((5, 7), 200, np.uint8)) == np.float32(200) * np.float32(1.0 / 255.0))


# ---- S5: octave plan closed form -------------------------------------------------------
def test_plan_closed_form():
    p = so.plan(912, 1368)
    assert p[0] == (1824, 2736) and p[-1] == (15, 22) and len(p) == 8
    p = so.plan(1978, 2640)
    assert p[0] == (3956, 5280) and p[-1] == (16, 21) and len(p) == 9
    assert len(so.plan(989, 1320)) == 8


# ---- golden vectors of the reference's CPU path ------------------------------------------
@pytest.mark.parametrize("name", ["tiny_cv", "c1_crop_cv", "c1_crop_pipe"])
def test_stages_vs_golden(name):
    g = golden(name)
    st = so.State(g["img"])
    gs = np.array([[float(np.sum(st.gauss(o, i).astype(np.float64))) for i in range(6)]
                   for o in range(len(st.octaves))])
    np.testing.assert_array_equal(gs, g["st_g_sums"])
    ds = np.array([float(np.sum(st.dog(o).astype(np.float64))) for o in range(len(st.octaves))])
    np.testing.assert_array_equal(ds, g["st_dog_sums"])
    if g["st_dog_oct0"].size:
        np.testing.assert_array_equal(st.dog(0), g["st_dog_oct0"])
    mag = np.concatenate([st.mag(o).ravel() for o in range(len(st.octaves))])[:4096]
    np.testing.assert_array_equal(mag, g["st_mag_first"])
    ang = np.concatenate([st.ang(o).ravel() for o in range(len(st.octaves))])[:4096]
    assert np.abs(ang - g["st_ang_first"]).max() <= 1e-4      # numpy's float32 atan2 ulps
    F = st.extrema(_contrast(g))
    np.testing.assert_array_equal(F, g["st_F"])
    o, l, r, c, X, resp, _ = st.refine(F, _contrast(g), _edge(g))
    for a, b in ((o, "o"), (l, "l"), (r, "r"), (c, "c"), (X, "X"), (resp, "resp")):
        np.testing.assert_array_equal(a, g["st_ref_" + b])
    scl = np.float32(1.6) * np.power(np.float32(2.0), (l.astype(np.float32) + X[:, 2]) / np.float32(3))
    for i in range(min(64, len(o))):
        h = st.ori_hist(o[i], l[i], r[i], c[i], g["st_scl"][i])
        np.testing.assert_allclose(h, g["st_hist"][i], rtol=2e-6, atol=1e-9)
    assert np.allclose(scl, g["st_scl"], rtol=1e-6)


def _contrast(g):
    # fixtures store float32; the reference ran with the Python floats 0.04 / 0.015
    return 0.04 if abs(float(g["contrast"]) - 0.04) < 1e-6 else 0.015


def _edge(g):
    return float(g["edge"])


@pytest.mark.parametrize("name", NAMES)
def test_detect_describe_vs_golden(name):
    g = golden(name)
    img = golden_image(name, g)
    k, st = so.detect(img, int(g["max_kp"]), _contrast(g), _edge(g))
    assert len(k["x"]) == len(g["kp_x"])
    for key in ("x", "y", "response", "octave", "layer"):
        np.testing.assert_array_equal(k[key], g["kp_" + key])
    np.testing.assert_allclose(k["size"], g["kp_size"], rtol=1e-6)
    # angle: 1e-3 relative (numpy's float32 atan2/exp ulps can move a sample across a bin)
    d = np.abs(k["angle"].astype(np.float64) - g["kp_angle"]) % 360
    assert (np.minimum(d, 360 - d) <= np.maximum(1e-3 * np.abs(g["kp_angle"]), 1e-3)).all()
    desc = so.describe(st, k, g["sel"])
    assert np.abs(desc - g["desc"]).max() <= 1e-2


# ---- S9: extrema == scipy maximum/minimum_filter formulation ------------------------------
def test_extrema_vs_scipy_filters():
    g = golden("c1_crop_pipe")
    st = so.State(g["img"])
    contrast = 0.015
    pthr = np.float32(0.8 * 0.5 * contrast / 3)
    masks = []
    for o in range(len(st.octaves)):
        D = st.dog(o)
        R, C = D.shape[1:]
        mx = ndi.maximum_filter(D, size=3, mode="nearest")
        mn = ndi.minimum_filter(D, size=3, mode="nearest")
        ext = ((D >= mx) & (D > pthr)) | ((D <= mn) & (D < -pthr))
        ext[0] = ext[-1] = False
        ext[:, :5] = ext[:, R - 5:] = False
        ext[:, :, :5] = ext[:, :, C - 5:] = False
        masks.append(ext.ravel())
    np.testing.assert_array_equal(st.extrema(contrast), np.nonzero(np.concatenate(masks))[0])


# ---- S12: orientation peaks closed forms ---------------------------------------------------
def test_ori_peaks_spike_and_parabola():
    h = np.zeros(36, np.float32)
    h[7] = 1.0
    np.testing.assert_array_equal(so.ori_peaks(h), np.float32([360 - 70]))
    h[0] = 1.0                        # second equal spike -> two peaks in bin order
    a = so.ori_peaks(h)
    assert list(a) == [0.0, np.float32(290.0)]
    # a smoothed parabola around bin 20 with vertex at 20.3
    j = np.arange(36)
    h = np.maximum(0, 100 - 3 * (j - 20.3) ** 2).astype(np.float32)
    a = so.ori_peaks(h)
    assert len(a) == 1 and abs(a[0] - (360 - 10 * 20.3)) < 0.5
    assert len(so.ori_peaks(np.zeros(36, np.float32))) == 0


# ---- S14/S15: descriptor invariants --------------------------------------------------------
def test_descriptor_invariants():
    g = golden("c1_crop_pipe")
    k, st = so.detect(g["img"], 12000, 0.015, 12.0)
    d = so.describe(st, k, np.arange(len(k["x"])))
    nrm = np.linalg.norm(d.astype(np.float64), axis=1)
    assert np.all(np.abs(nrm - 512) < 0.05)
    assert d.min() >= 0
    # clipping at 0.2 of the pre-clip norm bounds every bin by ~0.2 * 512 / (post-clip norm share)
    assert d.max() <= 512.0


def test_rootsift_closed_form():
    rng = np.random.default_rng(3)
    d = (rng.random((50, 128)) * 100).astype(np.float32)
    d[0] = 0
    u = so.rootsift_u8(d)
    ref = d / (np.abs(d).sum(1, keepdims=True) + np.float32(1e-7))
    ref = np.clip(np.sqrt(ref) * np.float32(512.0), 0, 255).astype(np.uint8)
    np.testing.assert_array_equal(u, ref)


def _ref_order(W, kh):
    """The reference's float64 order (centre, then pairs outer -> inner), numpy, unfused."""
    r2 = len(kh) - 1
    k = kh.astype(np.float64)
    acc = W[:, r2].astype(np.float64) * k[0]
    for j in range(r2, 0, -1):
        acc = acc + (W[:, r2 - j].astype(np.float64) + W[:, r2 + j].astype(np.float64)) * k[j]
    return acc


def _near_midpoint_windows(rng, n, kh):
    """Windows whose reference-order value lies within a few float64 ulps of a float32 rounding
    midpoint: random data, then two knobs (a coarse one at distance 1, a tiny one at the
    outermost tap) that steer the sum to a target at most 8 ulps from the midpoint above it."""
    r2 = len(kh) - 1
    k = kh.astype(np.float64)
    W = rng.random((n, 2 * r2 + 1), dtype=np.float32)
    W[:, r2 + 1] = 0.0
    W[:, 2 * r2] = 0.0
    S0 = _ref_order(W, kh)
    f = S0.astype(np.float32)
    f = np.where(f.astype(np.float64) > S0, np.nextafter(f, np.float32(-1)), f)      # largest f32 <= S0
    nf = np.nextafter(f, np.float32(2))
    m = (f.astype(np.float64) + nf.astype(np.float64)) / 2.0
    m = np.where(m <= S0, (nf.astype(np.float64) + np.nextafter(nf, np.float32(2)).astype(np.float64)) / 2.0, m)
    target = m + rng.integers(-8, 9, n) * np.spacing(m)
    c = ((target - S0) / k[1] * (1.0 - 2.0 ** -18)).astype(np.float32)
    W[:, r2 + 1] = c
    d1 = target - _ref_order(W, kh)
    keep = d1 > 0
    W, d1 = W[keep], d1[keep]
    W[:, 2 * r2] = (d1 / k[r2]).astype(np.float32)
    return W, _ref_order(W, kh)


@pytest.mark.parametrize("layer", [0, 2, 3, 4, 5])
def test_fused_blur_rounding_argument(layer):
    """The device evaluates each blur output with fused multiply-adds and accepts the result only
    when acc * (1 -+ 2^-46) round to the same float32 (else it re-evaluates in the reference's
    order).  On windows steered to within 8 float64 ulps of a float32 rounding midpoint the fused form
    alone disagrees with the reference (so the generator reaches the boundary), the checked
    form never does, and a margin far below the fused form's error bound (2^-60) does not
    suffice.  Ordinary image windows: zero disagreements and almost no fallbacks."""
    full = so.gauss_kernel(so.sigmas()[layer])
    r2 = len(full) // 2
    kh = full[r2:].copy()
    rng = np.random.default_rng(100 + layer)
    W, S = _near_midpoint_windows(rng, 20000, kh)
    assert len(W) > 19000
    assert np.all(np.isfinite(S)) and np.all(S > 0)
    bad_plain, _ = so.fused_blur_check(W, kh, 0.0, use_check=False)
    assert bad_plain >= 50, bad_plain
    bad, fb = so.fused_blur_check(W, kh, 2.0 ** -46, use_check=True)
    assert bad == 0 and fb > 0.99 * len(W), (bad, fb)
    bad_small, _ = so.fused_blur_check(W, kh, 2.0 ** -60, use_check=True)
    assert bad_small > 0
    # ordinary data: u8 / 255 windows and smooth windows
    U = (rng.integers(0, 256, (200000, 2 * r2 + 1)).astype(np.float32) * np.float32(1 / 255))
    bad_u, fb_u = so.fused_blur_check(U, kh, 2.0 ** -46, use_check=True)
    assert bad_u == 0 and fb_u < 50, fb_u
    G = rng.random((200000, 2 * r2 + 1), dtype=np.float32) ** 3
    bad_g, fb_g = so.fused_blur_check(G, kh, 2.0 ** -46, use_check=True)
    assert bad_g == 0 and fb_g < 50, fb_g


@pytest.mark.parametrize("layer", [0, 2, 3, 4, 5])
def test_fused_blur_midpoint_bits_test(layer):
    """The device's cheaper test: the fused value is taken unless its low 29 float64 mantissa
    bits lie within 256 ulps of a float32 rounding midpoint (or it is below 2^-120).  Near
    midpoints it never disagrees with the reference order and falls back almost always; with
    no margin (only the exact midpoint pattern caught) it does disagree, so the steered windows
    reach the boundary; ordinary image windows and exact zeros: no disagreement and almost no
    fallbacks."""
    full = so.gauss_kernel(so.sigmas()[layer])
    r2 = len(full) // 2
    kh = full[r2:].copy()
    rng = np.random.default_rng(200 + layer)
    W, S = _near_midpoint_windows(rng, 20000, kh)
    bad, fb = so.fused_blur_check(W, kh, 256, use_check=2)
    assert bad == 0 and fb > 0.99 * len(W), (bad, fb)
    bad_small, _ = so.fused_blur_check(W, kh, 0, use_check=2)
    assert bad_small > 0
    U = (rng.integers(0, 256, (200000, 2 * r2 + 1)).astype(np.float32) * np.float32(1 / 255))
    bad_u, fb_u = so.fused_blur_check(U, kh, 256, use_check=2)
    assert bad_u == 0 and fb_u < 50, fb_u
    G = rng.random((200000, 2 * r2 + 1), dtype=np.float32) ** 3
    bad_g, fb_g = so.fused_blur_check(G, kh, 256, use_check=2)
    assert bad_g == 0 and fb_g < 50, fb_g
    Z = np.zeros((1000, 2 * r2 + 1), np.float32)
    assert so.fused_blur_check(Z, kh, 256, use_check=2) == (0, 0)
    # results below 2^-120 (float32 subnormal territory) always take the reference order
    T = (rng.random((2000, 2 * r2 + 1)) * 2.0 ** -124).astype(np.float32)
    bad_t, fb_t = so.fused_blur_check(T, kh, 256, use_check=2)
    assert bad_t == 0 and fb_t == len(T), (bad_t, fb_t)


def test_midpoint_window_as_one_shift_add():
    """The device evaluates low29 - (0x10000000 - 256) <= 512 (32-bit unsigned) as
    (lo << 3) - (0x0FFFFF00 << 3) <= 4096: equal on every low word (exhaustive around the
    window's edges, random elsewhere)."""
    rng = np.random.default_rng(5)
    edges = np.concatenate([np.arange(-2000, 2000) + c for c in (0x0FFFFF00, 0x10000100, 0x1FFFFF00)]) % 2 ** 32
    lo = np.concatenate([edges, rng.integers(0, 2 ** 32, 2_000_000), [0, 2 ** 32 - 1, 0x0FFFFF00 | 0xE0000000]])
    lo = lo.astype(np.uint64)
    low29 = lo & 0x1FFFFFFF
    a = ((low29 - (0x10000000 - 256)) % 2 ** 32) <= 512
    b = (((lo << 3) % 2 ** 32 - (0x0FFFFF00 << 3)) % 2 ** 32) <= 4096
    assert np.array_equal(a, b) and a.sum() > 1000


def test_grouped_rounding_check_flags_exactly_the_groups_with_an_uncertain_output():
    """The blur folds each output's check quantities into two running unsigned minima, mid =
    (lo << 3) - (0x0FFFFF00 << 3) and tiny = bits(r) - 1 (32-bit), and re-evaluates the group
    output by output iff min(mid) <= 4096 or min(tiny) < 0x037FFFFF.  Groups of 16 and 14
    float64 results, some steered into the midpoint window or below 2^-120 (and exact zeros,
    which must not count as tiny): the group flag equals 'some output fails the per-output
    check'."""
    rng = np.random.default_rng(9)
    M = np.uint64(2 ** 32)
    for size in (16, 14):
        acc = rng.uniform(0.0, 1.0, (4000, size))
        bits = acc.view(np.uint64).copy()
        pick = rng.random(acc.shape) < 0.004                      # into the midpoint window
        bits[pick] = (bits[pick] & ~np.uint64(0x1FFFFFFF)) | np.uint64(0x10000000) + rng.integers(
            -300, 300, int(pick.sum())).astype(np.uint64)
        acc = bits.view(np.float64)
        tiny_pick = rng.random(acc.shape) < 0.003                 # below 2^-120, non-zero
        acc[tiny_pick] = rng.uniform(2.0 ** -140, 2.0 ** -121, int(tiny_pick.sum()))
        acc[rng.random(acc.shape) < 0.01] = 0.0
        lo = acc.view(np.uint64) & np.uint64(0xFFFFFFFF)
        r = acc.astype(np.float32)
        mid = ((lo << np.uint64(3)) % M - np.uint64(0x0FFFFF00 << 3)) % M
        tiny = (r.view(np.uint32).astype(np.uint64) - np.uint64(1)) % M
        per_output = (mid <= 4096) | (tiny < 0x037FFFFF)
        # the per-output check in its own terms: near a float32 midpoint, or 0 < r < 2^-120
        low29 = acc.view(np.uint64) & np.uint64(0x1FFFFFFF)
        near = ((low29 - np.uint64(0x10000000 - 256)) % M) <= 512
        assert np.array_equal(per_output, near | ((r > 0) & (r < np.float32(2.0 ** -120))))
        group = (mid.min(1) <= 4096) | (tiny.min(1) < 0x037FFFFF)
        assert np.array_equal(group, per_output.any(1))
        assert 0 < group.sum() < len(group)
