"""Drop-in replacement of the reference's SIFT module (app/planesite/gsift.py).

Same call signatures and array layouts as the reference:

* ``detect(gray, gpu=True, max_kp=12000, contrast=0.015, edge=12.0, timings=None)``
  -> ``(kps, state)`` with ``kps`` a dict of host arrays ``x, y, size, angle, response``
  (float32, input-image pixels / degrees) and ``octave, layer`` (int32, octave 0 = the 2x
  upsampled one), strongest first (gsift.py:435-477);
* ``describe(state, kps, sel, budget=2.5e6, timings=None)`` -> float32 ``(len(sel), 128)``
  descriptors with L2 norm 512, bin index ``(row*4 + col)*8 + orientation``
  (gsift.py:480-525).

Everything runs on the GPU through libfsx's C ABI; torch tensors are only device buffers.
``gpu=False`` is rejected: the CPU restatement lives in the test oracle, not in the product.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _lib

N_LAYERS = 3
SIGMA = 1.6
BORDER = 5
KEYS = ("x", "y", "size", "angle", "response", "octave", "layer")
# Module state the reference's GPU self-test touches (features.py:179-187 sets
# USE_CUDA_KERNELS[0] = False and clears _KERNELS when its CUDA kernels misbehave).  Here the
# kernels are the compiled libfsx, always used; the flag is ignored and the dict stays empty
# (there is no runtime-compiled kernel cache to clear).
USE_CUDA_KERNELS = [True]
_KERNELS = {}


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("fieldstitch_amd.gsift needs a ROCm GPU (torch.cuda.is_available() is False)")
    return torch


def plan(H: int, W: int):
    """Octave shapes (rows, cols) of the scale space of an H x W image (gsift.py:167-172)."""
    n = C.c_int()
    R = (C.c_int * 16)()
    Cc = (C.c_int * 16)()
    _lib.call("fsx_sift_plan", int(H), int(W), C.byref(n), R, Cc)
    return [(R[o], Cc[o]) for o in range(n.value)]


def taps(layer: int) -> np.ndarray:
    """float32 Gaussian taps the library uses (0 = base blur, 1..5 = layer increments)."""
    buf = (C.c_float * 64)()
    n = C.c_int()
    _lib.call("fsx_sift_taps", int(layer), buf, C.byref(n))
    return np.array(buf[:n.value], np.float32)


def _next_pow2(x: int) -> int:
    return 1 << max(0, int(x - 1).bit_length())


def default_caps(H: int, W: int, max_kp: int):
    """Candidate and keypoint capacities sized from the upsampled image area."""
    p0 = 4 * H * W
    cand = max(1 << 15, _next_pow2(p0 // 64))
    kp = max(_next_pow2(max(max_kp, 1024)), _next_pow2(cand))
    return cand, kp


class State:
    """Device workspace of one detected image: the Gaussian planes (layers 1..3 are what
    describe() samples; the DoG is their difference, never stored), plus the detect() output
    arrays (the reference's ``(xp, _Atlas)``)."""

    def __init__(self, shape, ws, kp_dev, n, max_kp, stream, stream_given=False):
        self.shape = shape
        self.ws = ws
        self.kp_dev = kp_dev
        self.n = n
        self.max_kp = max_kp
        self.stream = stream
        self.stream_given = stream_given


def _on_stream(stream):
    """Context that makes `stream` torch's current stream for the call (all allocations,
    launches and copies of the call are then ordered on it)."""
    from contextlib import nullcontext
    return _torch().cuda.stream(stream) if stream is not None else nullcontext()


def detect(gray, gpu=True, max_kp=12000, contrast=0.015, edge=12.0, timings=None, stream=None,
           caps=None):
    """Detect oriented keypoints (gsift.py:435-477).  Returns (dict of host arrays, state).
    gpu=False (the reference's CPU path) raises ValueError: this module has no CPU path.
    stream: every buffer, launch and copy of the call is ordered on it."""
    if not gpu:
        raise ValueError("fieldstitch_amd.gsift runs on the GPU only (gpu=False is the test oracle's job)")
    with _on_stream(stream):
        return _detect(gray, max_kp, contrast, edge, timings, stream, caps)


def _detect(gray, max_kp, contrast, edge, timings, stream, caps):
    torch = _torch()
    t0 = time.perf_counter()
    if isinstance(gray, torch.Tensor):
        img = gray if gray.is_cuda else gray.cuda()
        img = img.contiguous()
        H, W = img.shape
    else:
        g = np.ascontiguousarray(np.asarray(gray, np.uint8))
        H, W = g.shape
        img = torch.from_numpy(g).to("cuda", non_blocking=False)
    if img.dtype != torch.uint8:
        raise TypeError("gray must be uint8")
    max_kp = int(max_kp)
    cand_cap, kp_cap = caps if caps else default_caps(H, W, max_kp)
    st = _lib.stream_handle(stream)
    while True:
        nbytes = C.c_size_t()
        _lib.call("fsx_sift_workspace_bytes", H, W, cand_cap, kp_cap, C.byref(nbytes))
        ws = torch.empty(nbytes.value, dtype=torch.uint8, device=img.device)
        kp = torch.empty(7 * max_kp, dtype=torch.float32, device=img.device)
        n_dev = torch.zeros(1, dtype=torch.int32, device=img.device)
        stage = (C.c_float * 5)() if timings is not None else None
        _lib.call("fsx_sift_detect", C.c_void_p(img.data_ptr()), H, W, W, float(contrast), float(edge), max_kp,
                  cand_cap, kp_cap, C.c_void_p(ws.data_ptr()), nbytes.value, C.c_void_p(kp.data_ptr()),
                  C.c_void_p(n_dev.data_ptr()), stage, C.c_void_p(st))
        cnt = (C.c_int * 4)()
        _lib.call("fsx_sift_counters", C.c_void_p(ws.data_ptr()), cnt, C.c_void_p(st))
        if cnt[3] == 0:
            break
        # a capacity overflowed: grow and run again (the counters hold the needed sizes)
        cand_cap = _next_pow2(max(cand_cap * 2, cnt[0], cnt[1]))
        kp_cap = _next_pow2(max(kp_cap * 2, cnt[2], cand_cap))
    n = int(n_dev.item())
    host = kp.view(7, max_kp)[:, :n].cpu().numpy()
    kps = {k: host[i].copy() for i, k in enumerate(KEYS[:5])}
    kps["octave"] = host[5].view(np.int32).copy()
    kps["layer"] = host[6].view(np.int32).copy()
    if timings is not None:
        for name, v in zip(("pyramid", "extrema", "refine", "orientation", "select"), stage):
            timings[name] = timings.get(name, 0.0) + v / 1000.0
        timings["detect_wall"] = timings.get("detect_wall", 0.0) + time.perf_counter() - t0
    return kps, State((H, W), ws, kp, n, max_kp, st, stream is not None)


def describe(state: State, kps, sel, budget=2.5e6, timings=None, rootsift_u8=False):
    """128-d SIFT descriptors (L2 norm 512) for kps[sel] (gsift.py:480-525).  ``budget`` is
    accepted for signature compatibility (the GPU kernel needs no chunking).  With
    ``rootsift_u8=True`` returns (float32 desc, uint8 RootSIFT desc) (features.py:80-82).
    Runs on the stream the state was detected on."""
    with _on_stream(_torch().cuda.ExternalStream(state.stream) if state.stream_given else None):
        return _describe(state, kps, sel, timings, rootsift_u8)


def _describe(state, kps, sel, timings, rootsift_u8):
    torch = _torch()
    t0 = time.perf_counter()
    sel = np.asarray(sel, np.int64).reshape(-1)
    m = len(sel)
    if m == 0:
        d = np.zeros((0, 128), np.float32)
        return (d, np.zeros((0, 128), np.uint8)) if rootsift_u8 else d
    dev = state.ws.device
    n_all = len(kps["x"])
    if sel.min() < 0 or sel.max() >= n_all:
        raise IndexError("sel out of range")
    # the caller's keypoint arrays are authoritative (they may differ from detect()'s)
    kp = np.empty((7, n_all), np.float32)
    for i, k in enumerate(KEYS[:5]):
        kp[i] = np.asarray(kps[k], np.float32)
    kp[5] = np.asarray(kps["octave"], np.int32).view(np.float32)
    kp[6] = np.asarray(kps["layer"], np.int32).view(np.float32)
    n_oct = len(plan(*state.shape))
    if kp[5].view(np.int32).max() >= n_oct or kp[6].view(np.int32).min() < 1 or kp[6].view(np.int32).max() > 3:
        raise ValueError("octave/layer out of range for this state")
    kp_d = torch.from_numpy(kp).to(dev)
    sel_d = torch.from_numpy(sel).to(dev)
    out = torch.empty((m, 128), dtype=torch.float32, device=dev)
    u8 = torch.empty((m, 128), dtype=torch.uint8, device=dev) if rootsift_u8 else None
    _lib.call("fsx_sift_describe", C.c_void_p(state.ws.data_ptr()), C.c_void_p(kp_d.data_ptr()), n_all,
              C.c_void_p(sel_d.data_ptr()), m, C.c_void_p(out.data_ptr()),
              C.c_void_p(u8.data_ptr()) if u8 is not None else None, C.c_void_p(state.stream))
    d = out.cpu().numpy()
    if timings is not None:
        timings["describe"] = timings.get("describe", 0.0) + time.perf_counter() - t0
    if rootsift_u8:
        return d, u8.cpu().numpy()
    return d


def atan2_probe(n: int = 1 << 26, seed: int = 0x5EED) -> int:
    """Largest difference in float32 ulps between the device's atan2f and the float32 rounding
    of the float64 atan2 over n seeded gradient-like argument pairs (no reference
    counterpart: it bounds the orientation kernel's fast bin path)."""
    torch = _torch()
    out = torch.zeros(1, dtype=torch.int32, device="cuda")
    _lib.call("fsx_atan2_probe", C.c_longlong(int(n)), C.c_ulonglong(int(seed)), C.c_void_p(out.data_ptr()),
              C.c_void_p(_lib.stream_handle()))
    return int(out.item())


def desc_angle_probe(n: int = 1 << 26, seed: int = 0x5EED):
    """(descriptor angle error, atan2f-form error), both the largest circular difference in
    degrees from the float64 atan2 over n seeded gradient-like argument pairs (no reference
    counterpart: it bounds the descriptor's polynomial angle)."""
    torch = _torch()
    out = torch.zeros(2, dtype=torch.int32, device="cuda")
    _lib.call("fsx_desc_angle_probe", C.c_longlong(int(n)), C.c_ulonglong(int(seed)), C.c_void_p(out.data_ptr()),
              C.c_void_p(_lib.stream_handle()))
    e = out.cpu().numpy().view(np.uint32).astype(np.float64) * 1e-9
    return float(e[0]), float(e[1])
