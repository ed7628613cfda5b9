"""ctypes binding of ``libtrace_hip.so`` (the C-ABI in include/trace_hip.h).

Loading fails loudly if the library has not been built: there is no CPU
fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libtrace_hip.so")

TRACE_OK = 0
TRACE_E_INVALID = -1
TRACE_E_NODEVICE = -2
TRACE_E_HIP = -3
TRACE_E_OOM = -4
TRACE_E_UNSUPPORTED = -5
TRACE_E_DEVICE = -6

KEY_REF63 = 0
KEY_M30 = 1

MISS_DEFAULT, MISS_WHITE, MISS_SKY, MISS_SUN = 0, 1, 2, 3
MAT_DIFFUSE, MAT_LIGHT, MAT_REFLECTIVE, MAT_REFRACTIVE = 0, 1, 2, 3

_P = C.c_void_p
_I64 = C.c_int64
_I32 = C.c_int32
_DP = C.POINTER(C.c_double)
_FP = C.POINTER(C.c_float)
# trace_objective_fn: int (*)(void* user, const double* x, double* y)
OBJECTIVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _DP, _DP)


class TraceError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"trace error {code}: {msg}")
        self.code = code


class BvhInfo(C.Structure):
    _fields_ = [
        ("n_prims", C.c_int64),
        ("n_leaves", C.c_int64),
        ("n_inner", C.c_int64),
        ("depth", C.c_int32),
        ("key_mode", C.c_int32),
        ("arity", C.c_int32),
        ("pad", C.c_int32),
        ("root_box", C.c_double * 6),
        ("phr_initial_cut", C.c_int64),
        ("phr_levels", C.c_int32),
        ("n_segments", C.c_int32),
        ("phr_build_cost", C.c_int64),
    ]


# name -> argtypes; every entry returns int unless listed in _RESTYPES
SIGNATURES = {
    "trace_version": [],
    "trace_ctx_create": [C.c_int, C.POINTER(_P)],
    "trace_ctx_destroy": [_P],
    "trace_last_error": [_P],
    "trace_ctx_set_stream": [_P, _P],
    "trace_ctx_synchronize": [_P],
    "trace_mesh_create": [_P, _P, _I64, C.POINTER(_P)],
    "trace_mesh_create_device": [_P, _P, _I64, C.POINTER(_P)],
    "trace_mesh_update": [_P, _P, _P],
    "trace_mesh_destroy": [_P],
    "trace_lbvh_build": [_P, _P, _P, C.c_int, C.POINTER(_P)],
    "trace_lbvh_rebuild": [_P, _P, _P, C.c_int, _P],
    "trace_lbvh_rebuild_noalloc": [_P, _P, _P, C.c_int, _P],
    "trace_bvh_destroy": [_P],
    "trace_bvh_get_info": [_P, _P, C.POINTER(BvhInfo)],
    "trace_bvh_export_sorted": [_P, _P, _P, _P],
    "trace_bvh_export_canonical": [_P, _P, _I64, _P, _P, C.POINTER(_I64)],
    "trace_bvh_cost": [_P, _P, _DP],
    "trace_bvh_size": [_P, _P, C.POINTER(_I64)],
    "trace_phr_build": [_P, _P, C.c_double, C.c_double, C.c_int, _I32, C.POINTER(_P)],
    "trace_phr_rebuild": [_P, _P, C.c_double, C.c_double, C.c_int, _I32, _P],
    "trace_phr_export_source": [_P, _P, _P],
    "trace_closest_hit": [_P, _P, _P, _I64, _P, _P],
    "trace_closest_hit_device": [_P, _P, _P, _I64, _P, _P],
    "trace_instances_build": [_P, _P, C.c_int32, _P, _P, _P, C.c_int32, _P],
    "trace_bvh_device_bytes": [_P, _P, _P],
    "trace_closest_hit_counts": [_P, _P, _P, _I64, _P, _P],
    "trace_mesh_set_shading": [_P, _P, _P, _P, _P, _I32],
    "trace_render": [_P, _P, _P, _P, _P, _P, _P, _P],
    "trace_render_device": [_P, _P, _P, _P, _P, _P, _P, _P],
    "trace_mesh_create_segmented": [_P, _P, _I64, _P, _I32, C.POINTER(_P)],
    "trace_bvh_segment": [_P, _P, _I32, C.POINTER(_P)],
    "trace_render_batched": [_P, _P, _P, _I32, _P, _P, _P, _P, _P, _P],
    "trace_mesh_animate": [_P, _P, _P, C.c_double, C.c_double, C.c_uint64, _P],
    "trace_mesh_export": [_P, _P, _P],
    "trace_graph_begin": [_P],
    "trace_graph_end": [_P, C.POINTER(_P)],
    "trace_graph_launch": [_P, _P],
    "trace_graph_destroy": [_P],
    "trace_get_tune": [_P, _P],
    "trace_set_tune": [_P, _P],
    "trace_autotune_traversal": [_P, _P, _P, _I64, _P],
    "trace_autotune_render": [_P, _P, _P, _P, _P, _P],
    "trace_autotune_build": [_P, _P, C.c_int, _P],
    "trace_mesh_bind_params": [_P, _P, _P, _P],
    "trace_render_backward": [_P, _P, _P, _P, _P, _P, _P, _P],
    "trace_render_keep": [_P, _P, _P, _P, _P, _P, _P, _P, _P],
    "trace_render_backward_kept": [_P, _P, _P, _P, _P, _P, _P, _P, _P],
    "trace_render_segments": [_P, _P],
    "trace_render_counts": [_P, _P, _P, _P, _P, _P],
    "trace_mesh_set_shading_ex": [_P, _P, _P, _P, _P, _I32],
    "trace_scene_create": [_P, _P, _P, _P, _I32, _P, _P, _I32, _P],
    "trace_scene_compile": [_P, _P, _P, _P, _P, _I32, _P],
    "trace_scene_destroy": [_P],
    "trace_mesh_export_shading": [_P, _P, _P, _P],
    "trace_phr_build_with_cost": [_P, _P, C.c_double, C.c_double, C.c_int, _I32, _P, _P],
    "trace_phr_optimize_grid": [_P, _P, _P, _I32, _P, _I32, C.c_int, _I32, _I32, _I32, _P, _P, _P, _P],
    "trace_phr_optimize_bayes": [_P, _P, _P, _P, C.c_int, _I32, _I32, _I32, _I32, _I32, C.c_uint64, _P, _P, _P],
    "trace_bayes_minimize": [_P, _P, _I32, _I32, _I32, C.c_uint64, OBJECTIVE_FN, _P, _P, _P, _P],
    "trace_traversal_steps": [_P, _P, _P, _I64, _P],
    "trace_traversal_steps_device": [_P, _P, _P, _I64, _P],
    "trace_render_heatmap": [_P, _P, _P, _I32, _I32, _I32, _P, _P],
    "trace_render_heatmap_device": [_P, _P, _P, _I32, _I32, _I32, _P, _P],
    "trace_adam_step": [_P, _P, _P, _P, _P, _I64, C.c_float, C.c_float, C.c_float, C.c_float, _I32, C.c_float,
                        C.c_float],
}


class Tune(C.Structure):
    _fields_ = [
        ("trav_threads", C.c_int32),
        ("render_threads", C.c_int32),
        ("render_batch", C.c_int64),
        ("tuned", C.c_int32),
        ("trav_persistent", C.c_int32),
        ("render_persistent", C.c_int32),
        ("sort_bits", C.c_int32),
        ("sort_keys_per_thread", C.c_int32),
        ("trav_order", C.c_int32),
        ("render_order", C.c_int32),
        ("pad", C.c_int32),
    ]


class Camera(C.Structure):
    """trace_camera: origin, lower_left, horizontal, vertical."""

    _fields_ = [
        ("origin", C.c_double * 3),
        ("lower_left", C.c_double * 3),
        ("horizontal", C.c_double * 3),
        ("vertical", C.c_double * 3),
    ]


class RenderParams(C.Structure):
    _fields_ = [
        ("width", C.c_int32),
        ("height", C.c_int32),
        ("spp", C.c_int32),
        ("max_depth", C.c_int32),
        ("miss", C.c_int32),
        ("pad", C.c_int32),
        ("seed", C.c_uint64),
        ("sample_offset", C.c_int64),
    ]
_RESTYPES = {"trace_last_error": C.c_char_p}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (once).  Raises if it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"trace_amd: {LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )
    lib = C.CDLL(LIB_PATH)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, C.c_int)
    _lib = lib
    return lib


def check(rc: int, ctx=None) -> None:
    if rc != TRACE_OK:
        msg = ""
        if ctx is not None:
            m = load().trace_last_error(ctx)
            msg = m.decode() if m else ""
        raise TraceError(rc, msg)
