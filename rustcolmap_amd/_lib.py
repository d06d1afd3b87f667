"""ctypes binding of the C ABI in include/mvs_hip.h.

Loads the in-tree rustcolmap_amd/libmvs_hip.so.  There is no fallback: when
the library is missing or a call fails, an exception is raised.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmvs_hip.so")

MVS_OK = 0


class MvsOptionsC(C.Structure):
    _fields_ = [
        ("max_dim", C.c_uint32), ("num_sources", C.c_uint32), ("sweeps", C.c_uint32),
        ("num_hyp", C.c_uint32), ("min_consistent", C.c_uint32), ("save_depth_vis", C.c_uint32),
        ("depth_diff", C.c_double), ("voxel_rel", C.c_double), ("cost_thr", C.c_float),
        ("reserved", C.c_uint32),
    ]


class MvsViewC(C.Structure):
    _fields_ = [
        ("w", C.c_int32), ("h", C.c_int32), ("f", C.c_double), ("cx", C.c_double), ("cy", C.c_double),
        ("k", C.c_double), ("R", C.c_double * 9), ("t", C.c_double * 3),
        ("gray", C.POINTER(C.c_uint8)), ("rgb", C.POINTER(C.c_uint8)), ("name", C.c_char_p),
    ]


LOG_FN = C.CFUNCTYPE(None, C.c_char_p, C.c_void_p)   # mvs_log_fn


class MvsJobC(C.Structure):
    _fields_ = [
        ("ref", C.c_int32), ("k", C.c_int32), ("srcs", C.POINTER(C.c_int32)),
        ("zmin", C.c_double), ("zmax", C.c_double),
        ("depth", C.POINTER(C.c_float)), ("cost", C.POINTER(C.c_float)), ("normals", C.POINTER(C.c_float)),
    ]


class MvsPointC(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
                ("r", C.c_uint8), ("g", C.c_uint8), ("b", C.c_uint8), ("pad", C.c_uint8 * 5)]


class MvsStatsC(C.Structure):
    _fields_ = [
        ("n_eval_init", C.c_int64), ("n_eval_pass", C.c_int64), ("n_pixels", C.c_int64),
        ("taps_per_eval_sum", C.c_int64), ("ms_prepare", C.c_double), ("ms_init", C.c_double),
        ("ms_pass", C.c_double), ("ms_gate", C.c_double), ("ms_fuse", C.c_double),
        ("n_candidates", C.c_int64), ("n_points", C.c_int64), ("n_pass_launches", C.c_int32),
        ("graph_used", C.c_int32), ("ms_learned_volume", C.c_double), ("ms_learned_net", C.c_double),
        ("tile_w", C.c_int32), ("tile_h", C.c_int32), ("persistent", C.c_int32), ("evaluator", C.c_int32),
    ]


class MvsRunStatsC(C.Structure):
    _fields_ = [("views", C.c_int32), ("views_with_depth", C.c_int32), ("dense_points", C.c_int64),
                ("n_candidates", C.c_int64), ("cell", C.c_double), ("stereo_ms", C.c_double),
                ("fuse_ms", C.c_double), ("per_view_valid", C.POINTER(C.c_int64))]


class MvsMatchStatsC(C.Structure):
    _fields_ = [("pairs", C.c_int64), ("filtered_pairs", C.c_int64), ("distances", C.c_int64),
                ("matches", C.c_int64), ("overflow_items", C.c_int64), ("reserved_i", C.c_int64 * 3),
                ("exact_ms", C.c_double), ("tile_ms", C.c_double), ("collect_ms", C.c_double),
                ("recheck_ms", C.c_double), ("select_ms", C.c_double),
                ("host_launch_ms", C.c_double), ("host_wait_ms", C.c_double), ("host_out_ms", C.c_double),
                ("reserved", C.c_double * 4)]


# every symbol the header declares, with its ctypes signature
_P = C.c_void_p
_SIGS = {
    "mvs_options_default": (None, [C.POINTER(MvsOptionsC)]),
    "mvs_last_error": (C.c_char_p, []),
    "mvs_ctx_create": (C.c_int, [C.c_int, C.POINTER(_P)]),
    "mvs_ctx_destroy": (None, [_P]),
    "mvs_set_log": (C.c_int, [_P, LOG_FN, C.c_void_p]),
    "mvs_debug_rust_f64": (C.c_int, [C.c_double, C.c_char_p, C.c_int32]),
    "mvs_census": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P]),
    "mvs_census_noalloc": (C.c_int, [_P, _P, C.c_int32, C.c_int32, _P]),
    "mvs_upload_views": (C.c_int, [_P, C.POINTER(MvsViewC), C.c_int32]),
    "mvs_view_census": (C.c_int, [_P, C.c_int32, _P]),
    "mvs_estimate_depth": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.c_double, C.c_double,
                                     C.POINTER(MvsOptionsC), _P, _P, _P]),
    "mvs_estimate_depth_ex": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.c_double, C.c_double,
                                        C.POINTER(MvsOptionsC), C.c_int32, _P, _P, _P, _P, _P]),
    "mvs_estimate_depth_batch": (C.c_int, [_P, C.POINTER(MvsJobC), C.c_int32, C.POINTER(MvsOptionsC), C.c_int32]),
    "mvs_get_depth": (C.c_int, [_P, C.c_int32, _P, _P, _P]),
    "mvs_autotune": (C.c_int, [_P, C.POINTER(MvsJobC), C.c_int32, C.POINTER(MvsOptionsC), C.c_int32, _P, _P, _P,
                               _P]),
    "mvs_fuse": (C.c_int, [_P, _P, _P, _P, C.c_double, C.POINTER(MvsOptionsC),
                           C.POINTER(C.POINTER(MvsPointC)), C.POINTER(C.c_int64)]),
    "mvs_fuse_depth_maps": (C.c_int, [_P, _P, _P, _P, C.c_double, C.POINTER(MvsOptionsC),
                                      C.POINTER(C.POINTER(MvsPointC)), C.POINTER(C.c_int64)]),
    "mvs_free_points": (None, [C.POINTER(MvsPointC)]),
    "mvs_fusion_cell": (C.c_double, [_P, C.c_int64, C.c_double]),
    "mvs_fusion_diag": (C.c_double, [_P, C.c_int64]),
    "mvs_write_ply": (C.c_int, [C.c_char_p, _P, C.c_int64]),
    "mvs_write_dense_web_json": (C.c_int, [C.c_char_p, _P, C.c_int64]),
    "mvs_get_stats": (C.c_int, [_P, C.POINTER(MvsStatsC)]),
    "mvs_debug_plane_cost": (C.c_int, [_P, C.c_int32, _P, C.c_int32, _P, C.c_int32, _P, _P, _P, _P]),
    "mvs_debug_rng": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, _P, _P]),
    "mvs_debug_copy": (C.c_int, [_P, C.c_int64]),
    "mvs_debug_cvt": (C.c_int, [_P, _P, C.c_int32, _P]),
    "mvs_resize_image": (C.c_int, [_P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P]),
    "mvs_debug_div": (C.c_int, [_P, _P, _P, C.c_int32, _P]),
    "mvs_debug_run_passes": (C.c_int, [_P, C.c_int32, _P, C.c_int32, C.c_double, C.c_double, C.c_int32, C.c_int32,
                                       C.c_int32, _P, _P]),
    "mvs_plan": (C.c_int, [C.POINTER(MvsViewC), C.c_int32, _P, _P, _P, C.c_int64, C.POINTER(MvsOptionsC), _P, _P,
                          _P, _P, _P, _P]),
    "mvs_run_mvs": (C.c_int, [_P, C.POINTER(MvsViewC), C.c_int32, _P, _P, _P, C.c_int64, C.POINTER(MvsOptionsC),
                             C.c_char_p, C.POINTER(MvsRunStatsC)]),
    "mvs_net_create": (C.c_int, [_P, C.c_uint64, C.POINTER(_P)]),
    "mvs_net_destroy": (None, [_P]),
    "mvs_net_num_params": (C.c_int64, [_P]),
    "mvs_net_get_params": (C.c_int, [_P, _P]),
    "mvs_net_set_params": (C.c_int, [_P, _P]),
    "mvs_sweep_volume": (C.c_int, [_P, C.POINTER(MvsJobC), C.c_int32, _P]),
    "mvs_net_add_samples": (C.c_int, [_P, C.POINTER(MvsJobC), C.c_int32, _P]),
    "mvs_net_clear_samples": (C.c_int, [_P]),
    "mvs_net_train_step": (C.c_int, [_P, _P, C.c_int32, C.c_float, _P]),
    "mvs_net_predict": (C.c_int, [_P, C.POINTER(MvsJobC), _P]),
    "mvs_net_debug_grads": (C.c_int, [_P, _P, _P, C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P, _P,
                                      _P]),
    "mvs_estimate_depth_learned": (C.c_int, [_P, _P, C.POINTER(MvsJobC), C.c_int32, C.POINTER(MvsOptionsC),
                                             C.c_int32, C.c_float, C.POINTER(C.c_int64)]),
    "mvs_estimate_depth_learned_ex": (C.c_int, [_P, _P, C.POINTER(MvsJobC), C.c_int32, C.POINTER(MvsOptionsC),
                                                C.c_int32, C.c_float, C.c_int32, C.POINTER(C.c_int64)]),
    "mvs_match_pair": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, C.c_double, _P, _P, C.POINTER(C.c_int64)]),
    "mvs_upload_descriptors": (C.c_int, [_P, _P, _P, C.c_int32]),
    "mvs_match_frames": (C.c_int, [_P, _P, C.c_int32, C.c_double, _P, _P, _P]),
    "mvs_get_match_stats": (C.c_int, [_P, C.POINTER(MvsMatchStatsC)]),
    "mvs_match_set_mode": (C.c_int, [_P, C.c_int32]),
    "mvs_debug_match_bound": (C.c_int, [_P, _P, C.c_int32, _P, C.c_int32, _P]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises when it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "rustcolmap_amd: %s is missing; build it with `python -m rustcolmap_amd.build`" % LIB_PATH)
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class MvsError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str):
        super().__init__("%s failed (%d): %s" % (where, code, msg))
        self.code = code


def check(code: int, where: str):
    if code < MVS_OK:
        raise MvsError(code, where, load().mvs_last_error().decode("utf-8", "replace"))


def ptr(a):
    """Raw data pointer of a contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)
