"""ctypes declarations of the C ABI in include/blazert_amd.h (libblazert_amd.so, built in-tree).

The library is required: importing this module raises if it has not been built, so no caller can silently
fall back to a CPU path.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BLAZERT_AMD_LIB") or os.path.join(HERE, "libblazert_amd.so")

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_STATE = -2
ERR_DEVICE = -3
ERR_NO_DEVICE = -4
ERR_OUT_OF_MEMORY = -5
ERR_UNSUPPORTED = -6
INVALID_ID = 0xFFFFFFFF

GEOM_TRIANGLES, GEOM_SPHERES, GEOM_PLANES, GEOM_CYLINDERS = 0, 1, 2, 3
PRECISION_FP64 = 0
PRECISION_FP32_TRAV_FP64_HIT = 1
FLAG_HITS_INOUT = 1
FLAG_ANY_HIT = 2


class BuildOptions(C.Structure):
    _fields_ = [("min_leaf_primitives", C.c_uint32), ("max_tree_depth", C.c_uint32), ("bin_size", C.c_uint32)]


class BuildStats(C.Structure):
    _fields_ = [("max_tree_depth", C.c_uint32), ("leaf_nodes", C.c_uint32), ("max_num_prims_per_leaf", C.c_uint32),
                ("min_num_prims_per_leaf", C.c_uint32), ("branch_nodes", C.c_uint32), ("bad_splits", C.c_uint32),
                ("build_time_s", C.c_double), ("num_primitives", C.c_uint32), ("geom_id", C.c_uint32),
                ("device_bytes", C.c_uint64), ("built_on_device", C.c_uint32), ("reserved", C.c_uint32)]


class TreeNode(C.Structure):
    _fields_ = [("min", C.c_double * 3), ("max", C.c_double * 3), ("leaf", C.c_uint32), ("axis", C.c_uint32),
                ("children", C.c_uint32 * 2), ("prim_begin", C.c_uint32), ("prim_count", C.c_uint32)]


class Points(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p)]


class Rays(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("ox", "oy", "oz", "dx", "dy", "dz", "tmin", "tmax")]


class Hits(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "u", "v", "prim_id", "geom_id", "nx", "ny", "nz")]


SCHEDULE_PERSISTENT = 0
SCHEDULE_PER_THREAD = 1
SORT_NONE = 0
SORT_OCTANT_MORTON = 1
SORT_MORTON = 2
ARRAY_NODES64 = 0
ARRAY_NODES4 = 1
ARRAY_PRIMITIVES = 2
ARRAY_PRIM_IDS = 3


class Launch(C.Structure):
    _fields_ = [("precision", C.c_uint32), ("flags", C.c_uint32), ("block_size", C.c_uint32),
                ("schedule", C.c_uint32), ("refill", C.c_uint32), ("sort", C.c_uint32), ("stack_ring", C.c_uint32),
                ("pool_size", C.c_uint32)]


class BatchItem(C.Structure):
    _fields_ = [("scene", C.c_void_p), ("rays", Rays), ("hits", Hits), ("n", C.c_uint64)]


class Hit1(C.Structure):
    _fields_ = [("normal", C.c_double * 3), ("u", C.c_double), ("v", C.c_double), ("t", C.c_double),
                ("prim_id", C.c_uint32), ("geom_id", C.c_uint32)]


class Hit1F(C.Structure):  # brt_hit1_f32: RayHit<float>
    _fields_ = [("normal", C.c_float * 3), ("u", C.c_float), ("v", C.c_float), ("t", C.c_float),
                ("prim_id", C.c_uint32), ("geom_id", C.c_uint32)]


class MirrorOut(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("t", "u", "v", "prim_id", "px", "py", "pz", "rx", "ry", "rz")]


class MirrorGrads(C.Structure):
    _fields_ = [(n, C.c_void_p) for n in ("gt", "gpx", "gpy", "gpz", "vertices", "gox", "goy", "goz", "gdx", "gdy",
                                          "gdz")]


class TuneResult(C.Structure):
    _fields_ = [("launch", Launch), ("min_leaf_primitives", C.c_uint32), ("candidates", C.c_uint32),
                ("best_ms", C.c_double), ("default_ms", C.c_double)]


class BvhView(C.Structure):  # brt_bvh_view: device view of a user collection's tree
    _fields_ = [("nodes", C.c_void_p), ("prim_ids", C.c_void_p), ("root_min", C.c_double * 3),
                ("root_max", C.c_double * 3), ("root_ref", C.c_uint32), ("root_cnt", C.c_uint32),
                ("num_nodes", C.c_uint32), ("num_prims", C.c_uint32), ("depth", C.c_uint32),
                ("irregular_boxes", C.c_uint32), ("real_bytes", C.c_uint32), ("reserved", C.c_uint32)]


# every entry point declared in include/blazert_amd.h: name -> (restype, argtypes)
_P = C.c_void_p
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_F = C.POINTER(C.c_float)
SIGNATURES = {
    "brt_last_error": (C.c_char_p, []),
    "brt_version": (C.c_char_p, []),
    "brt_device_count": (C.c_int, []),
    "brt_scene_create": (C.c_int, [C.POINTER(BuildOptions), C.POINTER(C.c_void_p)]),
    "brt_scene_destroy": (C.c_int, [_P]),
    "brt_scene_add_mesh": (C.c_int, [_P, _D, C.c_uint32, _U, C.c_uint32, _U]),
    "brt_scene_create_f32": (C.c_int, [C.POINTER(BuildOptions), C.POINTER(C.c_void_p)]),
    "brt_scene_add_mesh_f32": (C.c_int, [_P, _F, C.c_uint32, _U, C.c_uint32, _U]),
    "brt_scene_add_spheres_f32": (C.c_int, [_P, _F, _F, C.c_uint32, _U]),
    "brt_scene_add_planes_f32": (C.c_int, [_P, _F, _F, _F, _F, C.c_uint32, _U]),
    "brt_scene_add_cylinders_f32": (C.c_int, [_P, _F, _F, _F, _F, _F, C.c_uint32, _U]),
    "brt_intersect_f32": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.POINTER(Launch),
                                    C.c_void_p]),
    "brt_intersect1_f32": (C.c_int, [_P, _F, _F, C.c_float, C.c_float, C.c_uint32, C.POINTER(Hit1F),
                                     C.POINTER(C.c_int)]),
    "brt_scene_add_spheres": (C.c_int, [_P, _D, _D, C.c_uint32, _U]),
    "brt_scene_add_planes": (C.c_int, [_P, _D, _D, _D, _D, C.c_uint32, _U]),
    "brt_scene_add_cylinders": (C.c_int, [_P, _D, _D, _D, _D, _D, C.c_uint32, _U]),
    "brt_scene_commit": (C.c_int, [_P]),
    "brt_scene_upload": (C.c_int, [_P]),
    "brt_scene_commit_device": (C.c_int, [_P, _P]),
    "brt_distance_to_surface": (C.c_int, [_P, C.c_uint32, C.POINTER(Points), _P, C.c_uint64, _P, _P]),
    "brt_closest_surface": (C.c_int, [_P, C.POINTER(Points), C.c_uint64, _P, _P, _P, _P]),
    "brt_scene_read_device_array": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint64,
                                              C.POINTER(C.c_uint64)]),
    "brt_scene_build_stats": (C.c_int, [_P, C.c_uint32, C.POINTER(BuildStats)]),
    "brt_scene_export_tree": (C.c_int, [_P, C.c_uint32, C.c_void_p, C.c_uint32, _U, _U]),
    "brt_intersect": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.POINTER(Launch), C.c_void_p]),
    "brt_intersect_noalloc": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.POINTER(Launch),
                                        C.c_void_p]),
    "brt_traverse": (C.c_int, [_P, C.c_uint32, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.POINTER(Launch),
                               C.c_void_p]),
    "brt_intersect1": (C.c_int, [_P, _D, _D, C.c_double, C.c_double, C.c_uint32, C.POINTER(Hit1),
                                 C.POINTER(C.c_int)]),
    "brt_batch_create": (C.c_int, [C.POINTER(BatchItem), C.c_uint32, C.POINTER(Launch), C.POINTER(C.c_void_p)]),
    "brt_batch_launch": (C.c_int, [_P, _P]),
    "brt_batch_capture": (C.c_int, [_P, _P]),
    "brt_batch_replay": (C.c_int, [_P, _P]),
    "brt_batch_destroy": (C.c_int, [_P]),
    "brt_autotune_launch": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.c_uint32, C.c_uint32,
                                      C.c_uint32, C.POINTER(TuneResult), C.c_void_p]),
    "brt_batch_autotune": (C.c_int, [_P, C.c_uint32, C.POINTER(TuneResult), C.c_void_p]),
    "brt_scene_tune_leaf": (C.c_int, [_P, C.POINTER(Rays), C.POINTER(Hits), C.c_uint64, C.c_uint32, _U, C.c_uint32,
                                      C.c_uint32, C.POINTER(TuneResult), C.c_void_p]),
    "brt_mirror_trace": (C.c_int, [_P, C.POINTER(Rays), C.c_uint64, C.c_uint32, C.c_double, C.POINTER(Launch),
                                   C.POINTER(MirrorOut), C.c_void_p]),
    "brt_mirror_adjoint": (C.c_int, [_P, C.POINTER(Rays), C.c_uint64, C.c_uint32, C.c_void_p, C.POINTER(MirrorGrads),
                                     C.c_void_p]),
    "brt_bvh_build": (C.c_int, [_D, _D, _D, C.c_uint32, C.POINTER(BuildOptions), C.POINTER(C.c_void_p)]),
    "brt_bvh_build_f32": (C.c_int, [_F, _F, _F, C.c_uint32, C.POINTER(BuildOptions), C.POINTER(C.c_void_p)]),
    "brt_bvh_device_view": (C.c_int, [_P, C.POINTER(BvhView)]),
    "brt_bvh_build_stats": (C.c_int, [_P, C.POINTER(BuildStats)]),
    "brt_bvh_export_tree": (C.c_int, [_P, C.c_void_p, C.c_uint32, _U, _U]),
    "brt_bvh_destroy": (C.c_int, [_P]),
}

_lib = None


class BlazertError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"blazert_amd error {code}: {message}")
        self.code = code


def lib():
    """The loaded C library (raises if libblazert_amd.so was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -c 'import __graft_entry__ as g; g.build()'` "
                              "or `python blazert_amd/build.py` first")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != OK:
        raise BlazertError(rc, lib().brt_last_error().decode())
