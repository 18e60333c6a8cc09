"""Fit of the odd polynomial that sift.hip's desc_angle_deg uses for atan on [0, 1].

atan(t) ~= t * P(t^2), P of degree 8 in t^2 (odd degree 17 in t).  Near-minimax fit by
iteratively reweighted least squares (Lawson's algorithm) on Chebyshev-distributed nodes,
weighting by the absolute error of atan.  Prints the coefficients (c0 .. c8, the constant
first; the kernel evaluates them in Horner form from c8 down) and the largest error of the
float32 Horner evaluation on a dense grid (about 1e-7 rad).

Run: python fieldstitch_amd/csrc/atan_poly_fit.py
"""
import numpy as np


def fit(deg: int = 8, nodes: int = 20000, iters: int = 200) -> np.ndarray:
    t = (1 - np.cos(np.linspace(0, np.pi, nodes))) / 2
    t = t[1:]
    f = np.arctan(t) / t
    A = np.stack([t ** (2 * k) for k in range(deg + 1)], 1)
    w = np.ones_like(t)
    c = None
    for _ in range(iters):
        sw = np.sqrt(w)
        c, *_ = np.linalg.lstsq(A * sw[:, None], f * sw, rcond=None)
        err = np.abs((A @ c - f) * t)
        w = w * (err / err.max() + 1e-3)
        w /= w.sum()
    return c


def horner_f32_error(c: np.ndarray, n: int = 2000001) -> float:
    c32 = c.astype(np.float32)
    tt = np.linspace(0, 1, n).astype(np.float32)
    z = (tt * tt).astype(np.float32)
    p = np.full_like(z, c32[-1])
    for k in range(len(c32) - 2, -1, -1):
        p = (p * z + c32[k]).astype(np.float32)
    r = (tt * p).astype(np.float32)
    return float(np.abs(r.astype(np.float64) - np.arctan(tt.astype(np.float64))).max())


if __name__ == "__main__":
    c = fit()
    print("coefficients c0..c8:", ", ".join("%.8g" % v for v in c))
    print("max |error| of the float32 Horner form on [0, 1]: %.3g rad" % horner_f32_error(c))
