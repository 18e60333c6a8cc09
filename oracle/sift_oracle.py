"""CPU oracle of the SIFT front end (test infrastructure only).

ctypes front end of ``sift_oracle.c``, the plain-C restatement of the reference's CPU path
(app/planesite/gsift.py).  The call shapes mirror the reference: ``detect(gray, max_kp,
contrast, edge) -> (kps, state)`` and ``describe(state, kps, sel) -> float32 (m, 128)``.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's CPU-baseline leg use this module, and
only as the checker or the CPU baseline.  Parity pins: tests/test_oracle_sift.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsift_oracle.so")

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with `make` (gcc, strict IEEE flags)."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "-s", "-C", HERE],
                              stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        _f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
        L.fso_fma.argtypes = [_f64p, _f64p, _f64p, _f64p, C.c_int64]
        L.fso_gauss_kernel.argtypes = [C.c_double, _f32p]
        L.fso_gauss_kernel.restype = C.c_int
        L.fso_sigmas.argtypes = [np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")]
        L.fso_plan.argtypes = [C.c_int, C.c_int, _i32p, _i32p]
        L.fso_plan.restype = C.c_int
        L.fso_blur.argtypes = [_f32p, _f32p, C.c_int, C.c_int, _f32p, C.c_int]
        L.fso_up2_u8.argtypes = [_u8p, C.c_int, C.c_int, C.c_int, _f32p]
        L.fso_sift_new.argtypes = [_u8p, C.c_int, C.c_int, C.c_int]
        L.fso_sift_new.restype = C.c_void_p
        L.fso_sift_free.argtypes = [C.c_void_p]
        L.fso_sift_n_oct.argtypes = [C.c_void_p]
        L.fso_sift_n_oct.restype = C.c_int
        L.fso_sift_octave.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
        for nm in ("fso_sift_dog", "fso_sift_mag", "fso_sift_ang"):
            getattr(L, nm).argtypes = [C.c_void_p, C.c_int]
            getattr(L, nm).restype = C.POINTER(C.c_float)
        L.fso_sift_gauss.argtypes = [C.c_void_p, C.c_int, C.c_int]
        L.fso_sift_gauss.restype = C.POINTER(C.c_float)
        L.fso_sift_extrema.argtypes = [C.c_void_p, C.c_double, _i64p, C.c_int64]
        L.fso_sift_extrema.restype = C.c_int64
        L.fso_sift_refine.argtypes = [C.c_void_p, _i64p, C.c_int64, C.c_double, C.c_double,
                                      _i32p, _i32p, _i32p, _i32p, _f32p, _f32p, _i64p]
        L.fso_sift_refine.restype = C.c_int64
        L.fso_sift_ori_hist.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, _f32p]
        L.fso_ori_peaks.argtypes = [_f32p, _f32p]
        L.fso_ori_peaks.restype = C.c_int
        L.fso_sift_detect.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int64,
                                      _f32p, _f32p, _f32p, _f32p, _f32p, _i32p, _i32p]
        L.fso_sift_detect.restype = C.c_int64
        L.fso_sift_describe.argtypes = [C.c_void_p, _f32p, _f32p, _f32p, _f32p, _i32p, _i32p,
                                        _i64p, C.c_int64, _f32p]
        L.fso_rootsift_u8.argtypes = [_f32p, C.c_int64, _u8p]
        L.fso_fused_blur_check.argtypes = [_f32p, C.c_int64, C.c_int, _f32p, C.c_double, C.c_int,
                                           C.POINTER(C.c_int64)]
        L.fso_fused_blur_check.restype = C.c_int64
        _lib = L
    return _lib


def fma(a, b, c) -> np.ndarray:
    """a * b + c with a single rounding (C99 fma), elementwise with numpy broadcasting."""
    a, b, c = np.broadcast_arrays(np.asarray(a, np.float64), np.asarray(b, np.float64), np.asarray(c, np.float64))
    shape = a.shape
    a, b, c = (np.ascontiguousarray(t).reshape(-1) for t in (a, b, c))
    out = np.empty(a.size, np.float64)
    if a.size:
        lib().fso_fma(a, b, c, out, a.size)
    return out.reshape(shape)


def gauss_kernel(sigma: float) -> np.ndarray:
    out = np.zeros(64, np.float32)
    k = lib().fso_gauss_kernel(float(sigma), out)
    return out[:k].copy()


def sigmas() -> np.ndarray:
    s = np.zeros(6, np.float64)
    lib().fso_sigmas(s)
    return s


def plan(H: int, W: int):
    R = np.zeros(16, np.int32)
    Cc = np.zeros(16, np.int32)
    n = lib().fso_plan(int(H), int(W), R, Cc)
    return [(int(R[o]), int(Cc[o])) for o in range(n)]


def blur(img: np.ndarray, k: np.ndarray) -> np.ndarray:
    img = np.ascontiguousarray(img, np.float32)
    out = np.empty_like(img)
    lib().fso_blur(img, out, img.shape[0], img.shape[1], np.ascontiguousarray(k, np.float32), len(k))
    return out


def fused_blur_check(windows: np.ndarray, half_taps: np.ndarray, margin: float, use_check=True):
    """Model of the device's fused blur evaluation against the reference order on (n, 2*r2+1)
    float32 windows: returns (windows whose float32 result differs, fallbacks taken).
    use_check: False = the fused value as is; True / 1 = accept when acc * (1 -+ margin) round
    alike; 2 = the device's test (margin = float64 ulps from a float32 rounding midpoint)."""
    w = np.ascontiguousarray(windows, np.float32)
    kh = np.ascontiguousarray(half_taps, np.float32)
    r2 = (w.shape[1] - 1) // 2
    assert w.shape[1] == 2 * r2 + 1 and len(kh) == r2 + 1
    fb = C.c_int64(0)
    bad = lib().fso_fused_blur_check(w.reshape(-1), w.shape[0], r2, kh, float(margin), int(use_check),
                                     C.byref(fb))
    return int(bad), int(fb.value)


def up2_u8(gray: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(gray, np.uint8)
    out = np.empty((2 * g.shape[0], 2 * g.shape[1]), np.float32)
    lib().fso_up2_u8(g, g.shape[0], g.shape[1], g.shape[1], out)
    return out


class State:
    """Gaussian/DoG pyramid and gradient atlas of one image (owner of the C allocation)."""

    def __init__(self, gray: np.ndarray):
        g = np.ascontiguousarray(gray, np.uint8)
        self.shape = g.shape
        self._p = lib().fso_sift_new(g, g.shape[0], g.shape[1], g.shape[1])
        n = lib().fso_sift_n_oct(self._p)
        self.octaves = []
        for o in range(n):
            R, Cc = C.c_int(), C.c_int()
            lib().fso_sift_octave(self._p, o, C.byref(R), C.byref(Cc))
            self.octaves.append((R.value, Cc.value))
        sz = [5 * r * c for r, c in self.octaves]
        self.dbase = np.r_[0, np.cumsum(sz)[:-1]].astype(np.int64)

    def __del__(self):
        if getattr(self, "_p", None):
            lib().fso_sift_free(self._p)
            self._p = None

    def _arr(self, ptr, n):
        return np.ctypeslib.as_array(ptr, shape=(n,)).copy()

    def dog(self, o: int) -> np.ndarray:
        R, Cc = self.octaves[o]
        return self._arr(lib().fso_sift_dog(self._p, o), 5 * R * Cc).reshape(5, R, Cc)

    def gauss(self, o: int, i: int) -> np.ndarray:
        R, Cc = self.octaves[o]
        return self._arr(lib().fso_sift_gauss(self._p, o, i), R * Cc).reshape(R, Cc)

    def mag(self, o: int) -> np.ndarray:
        R, Cc = self.octaves[o]
        return self._arr(lib().fso_sift_mag(self._p, o), 3 * R * Cc).reshape(3, R, Cc)

    def ang(self, o: int) -> np.ndarray:
        R, Cc = self.octaves[o]
        return self._arr(lib().fso_sift_ang(self._p, o), 3 * R * Cc).reshape(3, R, Cc)

    def extrema(self, contrast: float) -> np.ndarray:
        cap = 1 << 16
        while True:
            out = np.zeros(cap, np.int64)
            n = lib().fso_sift_extrema(self._p, float(contrast), out, cap)
            if n <= cap:
                return out[:n]
            cap = int(n)

    def refine(self, F: np.ndarray, contrast: float, edge: float):
        F = np.ascontiguousarray(F, np.int64)
        n = len(F)
        oo, ol, orr, oc = (np.zeros(n + 1, np.int32) for _ in range(4))
        X = np.zeros((n + 1) * 3, np.float32)
        resp = np.zeros(n + 1, np.float32)
        src = np.zeros(n + 1, np.int64)
        m = lib().fso_sift_refine(self._p, F, n, float(contrast), float(edge), oo, ol, orr, oc, X, resp, src)
        return oo[:m], ol[:m], orr[:m], oc[:m], X[:3 * m].reshape(m, 3), resp[:m], src[:m]

    def ori_hist(self, o: int, l: int, r: int, c: int, scl: float) -> np.ndarray:
        h = np.zeros(36, np.float32)
        lib().fso_sift_ori_hist(self._p, int(o), int(l), int(r), int(c), float(scl), h)
        return h


def ori_peaks(hist: np.ndarray) -> np.ndarray:
    a = np.zeros(36, np.float32)
    n = lib().fso_ori_peaks(np.ascontiguousarray(hist, np.float32), a)
    return a[:n].copy()


_KEYS = ("x", "y", "size", "angle", "response", "octave", "layer")


def detect(gray: np.ndarray, max_kp: int = 12000, contrast: float = 0.015, edge: float = 12.0):
    """gsift.detect (CPU semantics).  Returns (dict of host arrays, state)."""
    st = State(gray)
    cap = int(max_kp)
    f = [np.zeros(cap + 1, np.float32) for _ in range(5)]
    i = [np.zeros(cap + 1, np.int32) for _ in range(2)]
    n = lib().fso_sift_detect(st._p, float(contrast), float(edge), cap, *f, *i)
    vals = [a[:n].copy() for a in f + i]
    return dict(zip(_KEYS, vals)), st


def describe(state: State, kps: dict, sel) -> np.ndarray:
    """gsift.describe (CPU semantics): float32 (len(sel), 128), L2 norm 512."""
    sel = np.ascontiguousarray(np.asarray(sel, np.int64))
    out = np.zeros((len(sel), 128), np.float32)
    if len(sel) == 0:
        return out
    a = [np.ascontiguousarray(kps[k], np.float32) for k in ("x", "y", "size", "angle")]
    b = [np.ascontiguousarray(kps[k], np.int32) for k in ("octave", "layer")]
    lib().fso_sift_describe(state._p, *a, *b, sel, len(sel), out.reshape(-1))
    return out


def rootsift_u8(desc: np.ndarray) -> np.ndarray:
    """features._save RootSIFT encoding (float32 -> uint8)."""
    d = np.ascontiguousarray(desc, np.float32)
    out = np.zeros(d.shape, np.uint8)
    if len(d):
        lib().fso_rootsift_u8(d.reshape(-1), len(d), out.reshape(-1))
    return out
