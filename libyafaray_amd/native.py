"""ctypes binding of libymx.so (include/ymx_accel.h).

The shared library is the product; this module only loads it and describes
its structs.  It fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libymx.so"

YMX_OK = 0
YMX_NO_HIT = 0xFFFFFFFF
YMX_NO_VERTEX = 0xFFFFFFFF
YMX_VIS_VISIBLE = 1
YMX_VIS_CASTS_SHADOWS = 2


class YmxRay(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("tmin", C.c_float), ("dir", C.c_float * 3), ("tmax", C.c_float)]


class YmxHit(C.Structure):
    _fields_ = [("prim", C.c_uint32), ("t", C.c_float), ("u", C.c_float), ("v", C.c_float)]


class YmxAccelParams(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("max_leaf_size", C.c_int32), ("builder", C.c_int32),
                ("treelet_passes", C.c_int32), ("ray_reorder", C.c_int32), ("reorder_cell_bits", C.c_int32),
                ("reserved", C.c_int32 * 2)]


BUILDERS = {"lbvh": 0, "ploc": 1}


class YmxAccelInfo(C.Structure):
    _fields_ = [
        ("struct_size", C.c_uint32),
        ("n_tris", C.c_uint32),
        ("n_nodes", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("node_bytes", C.c_uint64),
        ("tri_bytes", C.c_uint64),
        ("build_ms", C.c_float),
        ("bound_min", C.c_float * 3),
        ("bound_max", C.c_float * 3),
        ("n_quads", C.c_uint32),
        ("treelet_passes", C.c_uint32),
        ("treelets_rewritten", C.c_uint32),
    ]


class YmxPtStats(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("camera_rays", C.c_uint64), ("extension_rays", C.c_uint64),
                ("shadow_rays", C.c_uint64), ("render_ms", C.c_float), ("guide_samples", C.c_uint64),
                ("guide_steps", C.c_uint64), ("light_trees", C.c_uint32)]


TUNING_TRIALS = 8


class YmxPtTuning(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("samples_per_wavefront", C.c_int32), ("shadow_kernel", C.c_int32),
                ("extension_kernel", C.c_int32), ("n_trials", C.c_int32),
                ("trial_samples_per_wavefront", C.c_int32 * TUNING_TRIALS),
                ("trial_shadow_kernel", C.c_int32 * TUNING_TRIALS),
                ("trial_extension_kernel", C.c_int32 * TUNING_TRIALS),
                ("trial_ms_per_sample", C.c_float * TUNING_TRIALS)]


# Every symbol include/ymx_accel.h and include/ymx_pt.h declare, with its ctypes signature.
SIGNATURES = {
    "ymx_pt_create": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_void_p,
                                C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p, C.c_uint32, C.c_void_p,
                                C.POINTER(C.c_void_p)]),
    "ymx_pt_destroy": (C.c_int, [C.c_void_p]),
    "ymx_pt_set_normals": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ymx_pt_clear_film": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ymx_pt_render": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p]),
    "ymx_pt_render_passes": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ymx_pt_read_film": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ymx_pt_get_stats": (C.c_int, [C.c_void_p, C.POINTER(YmxPtStats)]),
    "ymx_pt_set_guide": (C.c_int, [C.c_void_p, C.c_void_p, C.c_float, C.c_int32, C.c_uint32]),
    "ymx_pt_set_guide_depth": (C.c_int, [C.c_void_p, C.c_int32]),
    "ymx_pt_set_guide_selection": (C.c_int, [C.c_void_p, C.c_int32, C.c_float]),
    "ymx_pt_get_guide_selection": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "ymx_pt_set_guide_fallback": (C.c_int, [C.c_void_p, C.c_int32]),
    "ymx_pt_get_guide_fallback": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32]),
    "ymx_pt_autotune": (C.c_int, [C.c_void_p, C.c_int32, C.POINTER(YmxPtTuning), C.c_void_p]),
    "ymx_pt_get_tuning": (C.c_int, [C.c_void_p, C.POINTER(YmxPtTuning)]),
    "ymx_pt_set_tuning": (C.c_int, [C.c_void_p, C.POINTER(YmxPtTuning)]),
    "ymx_pt_save_film": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.c_int32, C.c_int32]),
    "ymx_pt_load_film": (C.c_int, [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                   C.POINTER(C.c_int32)]),
    "ymx_accel_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                   C.POINTER(YmxAccelParams), C.POINTER(C.c_void_p)]),
    "ymx_accel_create_faces": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                         C.POINTER(YmxAccelParams), C.POINTER(C.c_void_p)]),
    "ymx_accel_export_quads": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64]),
    "ymx_accel_destroy": (C.c_int, [C.c_void_p]),
    "ymx_accel_bound": (C.c_int, [C.c_void_p, C.c_float * 3, C.c_float * 3]),
    "ymx_accel_get_info": (C.c_int, [C.c_void_p, C.POINTER(YmxAccelInfo)]),
    "ymx_accel_intersect": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ymx_accel_intersect_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ymx_accel_intersect_shadow": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                             C.c_void_p]),
    "ymx_accel_export": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_uint64]),
    "ymx_accel_intersect_transparent": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p,
                                                  C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p]),
    "ymx_accel_traversal_stats": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                            C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.c_void_p]),
    "ymx_last_error": (C.c_char_p, []),
    # include/ymx_guide.h
    "ymx_guide_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "ymx_guide_destroy": (C.c_int, [C.c_void_p]),
    "ymx_guide_set_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ymx_guide_get_weights": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "ymx_guide_infer": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "ymx_guide_backward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "ymx_guide_get_gradients": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.POINTER(C.c_double)]),
    "ymx_guide_adam_step": (C.c_int, [C.c_void_p, C.c_void_p]),
    "ymx_guide_train_step": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
}


class YmxGuideParams(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("log2_table_size", C.c_uint32), ("n_min", C.c_int32),
                ("n_max", C.c_int32), ("bmin", C.c_float * 3), ("bmax", C.c_float * 3), ("max_batch", C.c_uint32),
                ("seed", C.c_uint32), ("lr", C.c_float), ("beta1", C.c_float), ("beta2", C.c_float),
                ("eps", C.c_float)]

_lib = None


class YmxError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libymx.so (built by __graft_entry__.build()); raise if absent."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise YmxError(f"{LIB_PATH} is not built; run __graft_entry__.build() first")
        # torch ships its own libamdhip64.so.7: load it first so that libymx.so
        # binds to the same HIP runtime (one runtime per process), whichever of
        # the two the caller touches first
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(rc: int, what: str) -> None:
    if rc != YMX_OK:
        msg = lib().ymx_last_error().decode(errors="replace")
        raise YmxError(f"{what} failed ({rc}): {msg}")
