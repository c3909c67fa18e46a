"""BVH over a user-defined collection (blazert::BVH<T, Collection>, blazert/bvh/accel.h:32-52).

The reference lets a caller plug in a primitive type through a collection class (examples/geometry_template/
GEOM_TEMPLATE.h:19-137) whose size(), get_primitive_bounding_box(i) and get_primitive_center(i) feed
SAHBinnedBuilder::build (blazert/bvh/builder.h:48-82). ``BVH`` builds that tree through the C ABI
(brt_bvh_build / brt_bvh_build_f32) from either such a collection object or plain (n, 3) arrays. Device code
traverses it with the caller's own intersector (include/blazert_amd/user_geometry.h); ``device_view()`` hands over
the tree's device pointers for that.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import capi
from .capi import check, lib

TREE_DTYPE = np.dtype([("min", np.float64, 3), ("max", np.float64, 3), ("leaf", np.uint32), ("axis", np.uint32),
                       ("children", np.uint32, 2), ("prim_begin", np.uint32), ("prim_count", np.uint32)])


def _gather(collection, dtype):
    n = int(collection.size())
    lo = np.zeros((n, 3), dtype)
    hi = np.zeros((n, 3), dtype)
    ctr = np.zeros((n, 3), dtype)
    for i in range(n):
        bmin, bmax = collection.get_primitive_bounding_box(i)
        lo[i], hi[i] = bmin, bmax
        ctr[i] = collection.get_primitive_center(i)
    return lo, hi, ctr


class BVH:
    """The reference's tree over a collection's boxes and centers.

    ``BVH(collection)`` with an object providing size() / get_primitive_bounding_box(i) / get_primitive_center(i),
    or ``BVH(box_min=..., box_max=..., centers=...)`` with (n, 3) arrays. dtype float64 builds BVH<double, ...>,
    float32 BVH<float, ...> (the builder's arithmetic in that type).
    """

    def __init__(self, collection=None, *, box_min=None, box_max=None, centers=None, dtype=np.float64,
                 min_leaf_primitives: int = 4, max_tree_depth: int = 28, bin_size: int = 64):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float64, np.float32):
            raise ValueError("dtype must be float64 or float32")
        if collection is not None:
            box_min, box_max, centers = _gather(collection, self.dtype)
        arrs = [np.ascontiguousarray(np.asarray(a, self.dtype).reshape(-1, 3)) for a in (box_min, box_max, centers)]
        n = len(arrs[0])
        if any(len(a) != n for a in arrs):
            raise ValueError("box_min, box_max and centers need the same number of rows")
        self.num_prims = n
        opts = capi.BuildOptions(min_leaf_primitives, max_tree_depth, bin_size)
        h = C.c_void_p()
        ptr = C.POINTER(C.c_double if self.dtype == np.float64 else C.c_float)
        fn = lib().brt_bvh_build if self.dtype == np.float64 else lib().brt_bvh_build_f32
        check(fn(*(a.ctypes.data_as(ptr) for a in arrs), n, C.byref(opts), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("BVH is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().brt_bvh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def build_stats(self) -> dict:
        s = capi.BuildStats()
        check(lib().brt_bvh_build_stats(self.handle, C.byref(s)))
        return {k: getattr(s, k) for k, _ in s._fields_}

    def export_tree(self):
        """Reference-form tree (pre-order nodes) and the leaf order of the primitives."""
        n = C.c_uint32(0)
        check(lib().brt_bvh_export_tree(self.handle, None, 0, None, C.byref(n)))
        nodes = np.zeros(n.value, dtype=TREE_DTYPE)
        order = np.zeros(max(self.num_prims, 1), dtype=np.uint32)
        check(lib().brt_bvh_export_tree(self.handle, nodes.ctypes.data_as(C.c_void_p), n.value,
                                        order.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(n)))
        return nodes, order[:self.num_prims]

    def device_view(self) -> capi.BvhView:
        """Uploads the tree to the current device (first call) and returns its brt_bvh_view."""
        v = capi.BvhView()
        check(lib().brt_bvh_device_view(self.handle, C.byref(v)))
        return v
