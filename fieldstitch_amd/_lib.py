"""ctypes binding of libfsx.so (the gfx950 HIP kernels behind a plain C ABI).

Every entry point returns an int status (0 = ok); a non-zero status becomes a RuntimeError
carrying ``fsx_last_error``, so callers keep the reference's exception-driven fallback logic.
There is no CPU fallback: if the library is missing, importing the product path fails.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfsx.so")

_vp = C.c_void_p
_i = C.c_int
_sz = C.c_size_t

# name -> argtypes (all return int status unless listed in _RESTYPE)
SIGNATURES = {
    "fsx_version": [],
    "fsx_last_error": [C.c_char_p, _sz],
    "fsx_device_info": [_i, C.c_char_p, _sz, C.POINTER(_i), C.POINTER(C.c_longlong)],
    "fsx_sift_plan": [_i, _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i)],
    "fsx_sift_taps": [_i, C.POINTER(C.c_float), C.POINTER(_i)],
    "fsx_sift_workspace_bytes": [_i, _i, _i, _i, C.POINTER(_sz)],
    "fsx_sift_detect": [_vp, _i, _i, _i, C.c_double, C.c_double, _i, _i, _i, _vp, _sz, _vp, _vp,
                        C.POINTER(C.c_float), _vp],
    "fsx_sift_counters": [_vp, C.POINTER(_i), _vp],
    "fsx_sift_describe": [_vp, _vp, _i, _vp, _i, _vp, _vp, _vp],
    "fsx_sift_table_bytes": [_i, C.POINTER(_sz)],
    "fsx_sift_features_batch": [_i, C.POINTER(_vp), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.c_double,
                                C.c_double, _i, _i, _i, _i, C.POINTER(_vp), C.POINTER(_sz), C.POINTER(_vp), _vp,
                                C.POINTER(_vp), C.POINTER(_vp), _vp, _sz, C.POINTER(_vp), _vp],
    "fsx_sift_features_prepare": [_i, C.POINTER(_vp), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.c_double,
                                  C.c_double, _i, _i, _i, _i, C.POINTER(_vp), C.POINTER(_sz), C.POINTER(_vp), _vp,
                                  C.POINTER(_vp), C.POINTER(_vp), _vp, _sz, _vp],
    "fsx_sift_features_launch": [_i, C.POINTER(_vp), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.c_double,
                                 C.c_double, _i, _i, _i, _i, C.POINTER(_vp), C.POINTER(_sz), C.POINTER(_vp), _vp,
                                 C.POINTER(_vp), C.POINTER(_vp), _vp, _sz, C.POINTER(_vp), _vp],
    "fsx_tune_set": [_i, _i],
    "fsx_tune_get": [_i, C.POINTER(_i)],
    "fsx_sift_selection_offset": [_i, _i, _i, _i, C.POINTER(_sz)],
    "fsx_match_workspace_bytes": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(_i), C.POINTER(_i),
                                  C.c_double, C.POINTER(_sz), C.POINTER(C.c_longlong)],
    "fsx_match_batch": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _vp, _vp, _i, C.POINTER(_i), C.POINTER(_i),
                        C.POINTER(_i), C.POINTER(C.c_double), C.c_double, _i, C.c_double, _i, _vp, _sz,
                        C.POINTER(C.c_longlong), _vp, _vp],
    "fsx_match_prepare": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _vp, _vp, _i, C.POINTER(_i),
                          C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_double), C.c_double, _i, C.c_double, _vp, _sz,
                          C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), _vp],
    "fsx_match_launch": [_vp, C.POINTER(C.c_longlong), _i, _vp, _vp],
    "fsx_match_batch_noalloc": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _vp, _vp, _i, C.POINTER(_i),
                                C.POINTER(_i), C.POINTER(_i), C.POINTER(C.c_double), C.c_double, _i, C.c_double, _i,
                                _vp, _vp, _vp, _vp, C.POINTER(C.c_longlong), _vp],
    "fsx_match_results": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(_i), C.POINTER(_i),
                          C.c_double, _vp, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _vp],
    "fsx_mfma_layout_probe": [_vp, _vp],
    "fsx_mfma_i8_layout_probe": [_vp, _vp],
    "fsx_atan2_probe": [C.c_longlong, C.c_ulonglong, _vp, _vp],
    "fsx_desc_angle_probe": [C.c_longlong, C.c_ulonglong, _vp, _vp],
    "fsx_undistort_table_bytes": [_i, C.POINTER(_sz)],
    "fsx_undistort_batch": [_i, C.POINTER(_i), C.POINTER(_vp), C.POINTER(C.c_double), C.POINTER(C.c_double), _i,
                            C.POINTER(_vp), C.POINTER(_vp), _vp, _sz, _vp],
    "fsx_undistort_f64": [_i, _vp, C.POINTER(C.c_double), _i, _vp, _vp],
    "fsx_ground_predict": [_i, _vp, C.POINTER(C.c_double), _vp, _vp, _vp],
    "fsx_pair_select": [_i, _vp, _vp, _i, _i, _vp, _vp],
    "fsx_clahe_workspace_bytes": [_i, _i, _i, C.POINTER(_sz)],
    "fsx_clahe": [_i, _vp, _i, _i, C.c_double, _i, _i, _vp, _vp, _sz, _vp],
    "fsx_tracks_workspace_bytes": [C.c_longlong, C.c_longlong, C.POINTER(_sz), C.POINTER(C.c_longlong)],
    "fsx_build_tracks": [_i, _vp, C.c_longlong, _i, _vp, _vp, _vp, _vp, C.c_longlong, _vp, _sz, C.c_longlong, _vp,
                         _vp, _vp, _vp, _vp],
    "fsx_fmatch_workspace_bytes": [_i, C.POINTER(_i), _i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_sz),
                                   C.POINTER(C.c_longlong)],
    "fsx_fmatch_batch": [_i, C.POINTER(_i), _vp, _i, C.POINTER(_i), C.POINTER(_i), C.c_double, _i, _i, _vp, _sz,
                         C.POINTER(C.c_longlong), _vp, _vp],
    "fsx_fmatch_results": [_i, C.POINTER(_i), _i, C.POINTER(_i), C.POINTER(_i), _vp, C.POINTER(_i), C.POINTER(_i),
                           C.POINTER(_i), C.POINTER(C.c_double), _vp],
    "fsx_fp8_layout_probe": [_vp, _vp],
    "fsx_fp8_scale_probe": [_vp, _vp],
    "fsx_fp8_cvt_probe": [_vp, _vp],
    "fsx_mlp_gemm": [_i, _i, _i, _vp, _i, _vp, _i, _i, _i, _i, _vp, _vp, _i, _vp, _vp, _vp],
    "fsx_mlp_splitk_scratch": [_i, _i, _i, C.POINTER(_sz)],
    "fsx_mlp_gemm_splitk": [_i, _i, _vp, _i, _vp, _i, _i, _i, _i, _vp, _i, _vp, _i, _vp],
    "fsx_mlp_head": [_vp, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp],
    "fsx_mlp_colsum_scratch": [_i, _i, C.POINTER(_sz)],
    "fsx_mlp_colsum": [_vp, _i, _i, _i, _i, _vp, _vp, _vp],
    "fsx_mlp_adam": [_vp, _vp, _vp, _vp, _vp, _i, C.c_float, C.c_float, C.c_float, C.c_float, _i, _vp],
    "fsx_mlp_to_bf16": [_vp, _vp, _i, _vp],
    "fsx_mlp_infer_pack_bytes": [C.POINTER(_sz)],
    "fsx_mlp_infer_pack": [_vp, _vp, _vp, _vp],
    "fsx_mlp_infer": [_vp, _i, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "fsx_match_candidate_offsets": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(_i),
                                    C.POINTER(_i), C.c_double, C.POINTER(C.c_longlong)],
    "fsx_ransac_homography": [_i, _vp, _vp, _vp, _i, C.c_double, C.c_ulonglong, _vp, _vp, _vp, _vp],
    "fsx_ransac_fundamental": [_i, _vp, _vp, _vp, _i, C.c_double, C.c_ulonglong, _vp, _vp, _vp, _vp],
    "fsx_ransac_similarity": [_i, _vp, _vp, _vp, _i, C.c_double, C.c_ulonglong, _vp, _vp, _vp, _vp],
    "fsx_match_filter_features": [_vp, C.POINTER(C.c_longlong), _vp, _vp, _vp, _vp],
    "fsx_match_contract": [_vp, C.POINTER(C.c_longlong), _i, _vp],
    "fsx_match_accepted": [_vp, C.POINTER(C.c_longlong), _vp, _vp, _vp],
    "fsx_match_filter_rows": [_vp, C.POINTER(C.c_longlong), _vp, _vp, _vp, _vp, C.c_longlong, _vp, _vp],
    "fsx_match_publish": [_vp, C.POINTER(C.c_longlong), _vp, _vp, C.c_longlong, _vp, C.c_float, _vp],
    "fsx_match_pack": [_vp, C.POINTER(C.c_longlong), _vp, C.c_longlong, _vp, _vp, _vp],
    "fsx_ransac_batch": [_i, _i, _vp, _vp, _vp, _i, _vp, _vp, _vp, _vp, _vp, _vp],
    "fsx_ransac_batch_adaptive": [_i, _i, _vp, _vp, _vp, _i, C.c_double, _vp, _vp, _vp, _vp, _vp, _vp, _vp],
    "fsx_gather_xy": [C.c_longlong, _vp, _vp, C.c_longlong, _vp, _vp],
    "fsx_gather_lists": [_i, _vp, _vp, _vp, C.c_longlong, _vp, _vp, _vp, _vp, _vp, _vp, C.c_longlong, _vp, _vp, _vp,
                         _vp, _vp, _vp],
    "fsx_similarity_fit_batch": [_i, _vp, _vp, _vp, _vp, _vp, _vp],
    "fsx_prior_keep_batch": [_i, _vp, _vp, _vp, _vp, _vp, _i, _vp, _vp],
    "fsx_match_query_state": [_i, C.POINTER(_i), C.POINTER(_i), C.POINTER(_i), _i, C.POINTER(_i), C.POINTER(_i),
                              C.c_double, _vp, C.POINTER(_i), C.POINTER(C.c_double), C.POINTER(_i), _vp],
}

_lib = None


class FsxError(RuntimeError):
    pass


def load(path: str = LIB_PATH):
    """Load libfsx.so (raises OSError if it has not been built)."""
    global _lib
    if _lib is None:
        # torch first: its HIP runtime must be the one libfsx binds to (one runtime per process)
        import torch  # noqa: F401
        if not os.path.exists(path):
            raise OSError("libfsx.so not built (%s); run fieldstitch_amd/build.py or __graft_entry__.build()" % path)
        L = C.CDLL(path)
        for name, args in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = C.c_int
        _lib = L
    return _lib


def last_error() -> str:
    L = load()
    buf = C.create_string_buffer(1024)
    L.fsx_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def call(name: str, *args) -> int:
    rc = getattr(load(), name)(*args)
    if rc != 0:
        raise FsxError("%s failed (status %d): %s" % (name, rc, last_error()))
    return rc


def stream_handle(stream=None) -> int:
    """Raw hipStream_t of a torch stream (default: the current stream)."""
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return int(s.cuda_stream)
