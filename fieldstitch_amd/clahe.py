"""Contrast-limited adaptive histogram equalisation on the GPU: the conditioning every photo
gets before feature extraction (features.py:60-71, ``cv2.createCLAHE(clipLimit=2.0,
tileGridSize=(8, 8)).apply(gray)``; also :197 and :225).

``createCLAHE(clipLimit, tileGridSize).apply(gray)`` mirrors OpenCV's object: uint8 (H, W)
numpy in, numpy out; a uint8 CUDA tensor of (H, W) or a batch (n, H, W) in, a tensor on the
same device out.  The kernels (libfsx ``fsx_clahe``) follow OpenCV's 8-bit algorithm step by
step; with OpenCV absent here, equality with cv2's own output is not tested.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib


class CLAHE:
    def __init__(self, clipLimit=2.0, tileGridSize=(8, 8)):
        self.clip = float(clipLimit)
        self.tiles_x, self.tiles_y = (int(tileGridSize[0]), int(tileGridSize[1]))
        self._ws = None

    def getClipLimit(self):
        return self.clip

    def getTilesGridSize(self):
        return (self.tiles_x, self.tiles_y)

    def apply(self, gray, out=None, stream=None):
        """Equalise gray (see the module docstring).  out (optional, device input only): a
        contiguous uint8 CUDA tensor of the input's shape on its device, written in place.
        stream: all of the call's work (buffers, kernels, copies) is ordered on it."""
        import torch
        from contextlib import nullcontext
        with torch.cuda.stream(stream) if stream is not None else nullcontext():
            return self._apply(gray, out, stream)

    def _apply(self, gray, out, stream):
        import torch
        host = isinstance(gray, np.ndarray)
        if host:
            if gray.dtype != np.uint8 or gray.ndim != 2:
                raise TypeError("CLAHE.apply: expected a uint8 (H, W) image")
            if out is not None:
                raise TypeError("CLAHE.apply: out is only accepted with a device input")
            src = torch.from_numpy(np.ascontiguousarray(gray)).cuda()
        else:
            if gray.dtype != torch.uint8 or not gray.is_cuda or gray.ndim not in (2, 3):
                raise TypeError("CLAHE.apply: expected a uint8 CUDA tensor of (H, W) or (n, H, W)")
            src = gray.contiguous()
        batch = src.reshape((-1,) + tuple(src.shape[-2:]))
        n, H, W = batch.shape
        nb = C.c_size_t()
        _lib.call("fsx_clahe_workspace_bytes", n, self.tiles_x, self.tiles_y, C.byref(nb))
        if self._ws is None or self._ws.numel() < nb.value or self._ws.device != batch.device:
            self._ws = torch.empty(max(1, nb.value), dtype=torch.uint8, device=batch.device)
        if out is not None:
            if (not isinstance(out, torch.Tensor) or out.dtype != torch.uint8 or not out.is_cuda
                    or out.device != batch.device or tuple(out.shape) != tuple(src.shape) or not out.is_contiguous()):
                raise ValueError("CLAHE.apply: out must be a contiguous uint8 CUDA tensor of the input's shape "
                                 "on its device")
        dst = torch.empty_like(batch) if out is None else out.view(batch.shape)
        _lib.call("fsx_clahe", n, C.c_void_p(batch.data_ptr()), H, W, C.c_double(self.clip), self.tiles_x,
                  self.tiles_y, C.c_void_p(dst.data_ptr()), C.c_void_p(self._ws.data_ptr()), nb.value,
                  C.c_void_p(_lib.stream_handle(stream)))
        dst = dst.reshape(src.shape)
        return dst.cpu().numpy() if host else dst


def createCLAHE(clipLimit=2.0, tileGridSize=(8, 8)):
    """cv2.createCLAHE's signature; the returned object's apply runs on the GPU."""
    return CLAHE(clipLimit, tileGridSize)
