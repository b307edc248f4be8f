"""ctypes binding of liblauematch.so (the C ABI declared in include/lauematch.h).

The library is the only compute path: if it is missing or fails to load, every
call raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "liblauematch.so")

LM_OK = 0
LM_ERR = {-1: "LM_ERR_INVALID", -2: "LM_ERR_HIP", -3: "LM_ERR_NOMEM", -4: "LM_ERR_BUDGET",
          -5: "LM_ERR_OVERFLOW", -6: "LM_ERR_STATE"}
LM_FWD_FP64 = 0
LM_FWD_FP32_GUARD = 1

c_double_p = ctypes.POINTER(ctypes.c_double)


class LmGeom(ctypes.Structure):
    _fields_ = [("recip", (ctypes.c_double * 3) * 3), ("rotT", (ctypes.c_double * 3) * 3),
                ("P", ctypes.c_double * 3), ("pxX", ctypes.c_double), ("pxY", ctypes.c_double),
                ("nrPxX", ctypes.c_int32), ("nrPxY", ctypes.c_int32),
                ("Elo", ctypes.c_double), ("Ehi", ctypes.c_double)]


class LmFwdStats(ctypes.Structure):
    _fields_ = [("n_orients", ctypes.c_uint64), ("n_spots", ctypes.c_uint64),
                ("n_pairs", ctypes.c_uint64), ("n_recomputed", ctypes.c_uint64),
                ("ms", ctypes.c_double), ("n_projected", ctypes.c_uint64)]


class LmHits(ctypes.Structure):
    _fields_ = [("nr_results", ctypes.c_uint64), ("count", ctypes.c_uint64),
                ("ids", ctypes.POINTER(ctypes.c_uint64)), ("nspots", ctypes.POINTER(ctypes.c_int32)),
                ("totint", c_double_p), ("score", c_double_p), ("kernel_ms", ctypes.c_double),
                ("walk_ms", ctypes.c_double)]


class LaueMatchError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{LM_ERR.get(code, code)}: {msg}")
        self.code = code


class LmPrepConfig(ctypes.Structure):
    _fields_ = [("threshold_method", ctypes.c_int32), ("threshold_value", ctypes.c_double),
                ("threshold_percentile", ctypes.c_double), ("min_area", ctypes.c_int32),
                ("px_x", ctypes.c_double), ("distance", ctypes.c_double),
                ("orientation_spacing", ctypes.c_double), ("filter_radius", ctypes.c_int32),
                ("median_passes", ctypes.c_int32)]


class LmPrepResult(ctypes.Structure):
    _fields_ = [("threshold", ctypes.c_double), ("sigma", ctypes.c_double),
                ("n_components", ctypes.c_int32), ("n_kept", ctypes.c_int32),
                ("im_scale", ctypes.c_double), ("ms", ctypes.c_double),
                ("n_above", ctypes.c_uint64)]


class LmServerConfig(ctypes.Structure):
    _fields_ = [("max_batch", ctypes.c_int32), ("workers", ctypes.c_int32),
                ("max_inflight", ctypes.c_int32), ("use_graphs", ctypes.c_int32),
                ("batch_after_ms", ctypes.c_int32)]


class LmServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double),
                ("images_per_s", ctypes.c_double), ("graph_launches", ctypes.c_uint64)]


class LmTuneReport(ctypes.Structure):
    _fields_ = [("n_widths", ctypes.c_int32), ("widths", ctypes.c_int32 * 7),
                ("ms_single_table", ctypes.c_double), ("ms_single", (ctypes.c_double * 8) * 4),
                ("ms_lane", ctypes.c_double * 7), ("ms_grouped", ctypes.c_double * 7),
                ("ms_banded", (ctypes.c_double * 5) * 7), ("ms_nibble", ctypes.c_double * 7),
                ("banded_tile_bits", ctypes.c_int32), ("banded_loads", ctypes.c_int32),
                ("batch_tile_bits", ctypes.c_int32), ("batch_widths_mask", ctypes.c_uint32),
                ("grouped_mask", ctypes.c_uint32), ("nibble_mask", ctypes.c_uint32),
                ("batch_tile_bits_w", ctypes.c_int32 * 7),
                ("ms_banded16", (ctypes.c_double * 5) * 7), ("sum16_mask", ctypes.c_uint32),
                ("split_mask", ctypes.c_uint32), ("banded_split", ctypes.c_int32),
                ("ms_split", ctypes.c_double * 4)]


_lib = None

# name -> (restype, argtypes); every symbol declared in include/lauematch.h
SIGNATURES = {
    "lm_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint64, ctypes.POINTER(ctypes.c_void_p)]),
    "lm_ctx_destroy": (None, [ctypes.c_void_p]),
    "lm_ctx_bytes_in_use": (ctypes.c_uint64, [ctypes.c_void_p]),
    "lm_last_error": (ctypes.c_char_p, []),
    "lm_version": (ctypes.c_char_p, []),
    "lm_geom_init": (ctypes.c_int, [c_double_p, ctypes.c_int32, c_double_p, c_double_p,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                    ctypes.POINTER(LmGeom)]),
    "lm_quantize_u8": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                      c_double_p]),
    "lm_db_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                    ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(LmGeom),
                                    ctypes.c_int32, ctypes.POINTER(ctypes.c_void_p)]),
    "lm_db_destroy": (None, [ctypes.c_void_p]),
    "lm_db_size": (ctypes.c_uint64, [ctypes.c_void_p]),
    "lm_db_table_bytes": (ctypes.c_uint64, [ctypes.c_void_p]),
    "lm_db_table_spots": (ctypes.c_uint64, [ctypes.c_void_p]),
    "lm_db_forward": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.POINTER(LmFwdStats)]),
    "lm_db_table_checksum": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_uint64)]),
    "lm_db_set_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64]),
    "lm_db_get_option": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                        ctypes.POINTER(ctypes.c_int64)]),
    "lm_db_autotune": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(LmTuneReport)]),
    "lm_db_import_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64]),
    "lm_db_export_table": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                          ctypes.c_void_p]),
    "lm_score": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int32,
                                ctypes.c_double, ctypes.POINTER(LmHits)]),
    "lm_score_f64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                    ctypes.c_double, ctypes.POINTER(LmHits)]),
    "lm_score_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                    ctypes.c_int32, ctypes.c_double, ctypes.c_void_p,
                                    ctypes.POINTER(LmHits)]),
    "lm_quantize_u8_f32": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                          c_double_p]),
    "lm_score_dev_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double,
                                            ctypes.c_int32, ctypes.c_double, ctypes.c_void_p,
                                            ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.POINTER(ctypes.c_uint64),
                                            ctypes.POINTER(ctypes.c_uint64)]),
    "lm_score_batch": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                      ctypes.POINTER(ctypes.c_void_p), c_double_p, ctypes.c_int32,
                                      ctypes.c_double, ctypes.POINTER(LmHits)]),
    "lm_score_batch_each": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32,
                                           ctypes.POINTER(ctypes.c_void_p), c_double_p,
                                           ctypes.POINTER(ctypes.c_int32), c_double_p,
                                           ctypes.POINTER(LmHits)]),
    "lm_score_batch_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                          c_double_p, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_void_p, ctypes.POINTER(LmHits)]),
    "lm_score_fwd_batch_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                              c_double_p, ctypes.POINTER(ctypes.c_int32),
                                              c_double_p, ctypes.c_int32, ctypes.c_void_p,
                                              ctypes.POINTER(LmHits),
                                              ctypes.POINTER(LmFwdStats)]),
    "lm_hits_free": (None, [ctypes.POINTER(LmHits)]),
    "lm_preprocess": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                     ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(LmPrepConfig), ctypes.c_void_p,
                                     ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.POINTER(LmPrepResult)]),
    "lm_median_background": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                            ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]),
    "lm_merge_candidates_dev": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_uint64,
                                               ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32,
                                               ctypes.c_double, ctypes.c_int32, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.POINTER(ctypes.c_int32)]),
    "lm_merge_candidates": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_int32, ctypes.c_double,
                                           ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                           ctypes.c_void_p, ctypes.c_void_p,
                                           ctypes.POINTER(ctypes.c_int32)]),
    "lm_quat_from_matrix": (None, [ctypes.c_void_p, ctypes.c_void_p]),
    "lm_misorientation_deg": (ctypes.c_double, [ctypes.c_void_p, ctypes.c_void_p,
                                                ctypes.c_int32]),
    "lm_server_create": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LmServerConfig),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "lm_server_add_db": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "lm_server_submit": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p,
                                        ctypes.c_double, ctypes.c_int32, ctypes.c_double,
                                        ctypes.POINTER(ctypes.c_uint64)]),
    "lm_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(LmHits),
                                      ctypes.POINTER(ctypes.c_double)]),
    "lm_server_poll": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.POINTER(LmHits),
                                      ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_int32)]),
    "lm_server_get_stats": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(LmServerStats)]),
    "lm_server_reset_stats": (None, [ctypes.c_void_p]),
    "lm_server_destroy": (None, [ctypes.c_void_p]),
    "lm_selftest_fp64": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]),
}


def lib():
    """Load liblauematch.so (raises if it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run `python -m lauematching_amd.build`")
        # One HIP runtime per process: PyTorch's ROCm build bundles its own
        # libamdhip64, and two runtimes loaded side by side cannot both open
        # the device (whichever initialises second finds no GPU).  Importing
        # torch first makes this library bind to the runtime already loaded.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != LM_OK:
        raise LaueMatchError(rc, lib().lm_last_error().decode(errors="replace"))
    return rc
