"""User-defined collections, host side: the tree brt_bvh_build / brt_bvh_build_f32 builds over a caller's boxes and
centers (blazert::BVH<T, Collection> + SAHBinnedBuilder::build, blazert/bvh/builder.h:48-82) against the oracle's
builder on the same arrays, plus the collection-object path, the C++ template path of the example plugin and the
argument checks. No GPU needed.
"""
import ctypes as C
import os

import numpy as np
import pytest

from blazert_amd import BVH, BlazertError, Scene
from blazert_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "examples", "user_geometry", "libuser_geometry.so")
FIELDS = ("min", "max", "leaf", "axis", "children", "prim_begin", "prim_count")


def _boxes(n, seed):
    """Random axis-aligned boxes of mixed aspect (not the cubes of a sphere collection) with off-center centers."""
    r = W.Rng(seed)
    c = r.uniform(3 * n, -20.0, 20.0).reshape(n, 3)
    half = np.exp(r.uniform(3 * n, -3.0, 1.0)).reshape(n, 3)
    lo, hi = c - half, c + 0.5 * half
    ctr = lo + r.uniform(3 * n, 0.2, 0.8).reshape(n, 3) * (hi - lo)
    return lo, hi, ctr


def _same_tree(a, b):
    for f in FIELDS:
        x, y = np.asarray(a[f]), np.asarray(b[f])
        if x.dtype.kind == "f":
            assert np.array_equal(x.astype(np.float64).view(np.uint64), y.astype(np.float64).view(np.uint64)), f
        else:
            assert np.array_equal(x, y), f


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("opts", [(4, 28, 64), (1, 28, 16), (8, 6, 8), (2, 40, 2)])
def test_collection_tree_matches_oracle(oracle_mod, dtype, opts):
    lo, hi, ctr = _boxes(6000, 11)
    b = BVH(box_min=lo, box_max=hi, centers=ctr, dtype=dtype, min_leaf_primitives=opts[0], max_tree_depth=opts[1],
            bin_size=opts[2])
    nodes, order = b.export_tree()
    ref_nodes, ref_order, ref_stats = oracle_mod.collection_tree(lo, hi, ctr, dtype, *opts)
    _same_tree(nodes, ref_nodes)
    assert np.array_equal(order, ref_order)
    st = b.build_stats()
    for k in ("max_tree_depth", "leaf_nodes", "branch_nodes", "bad_splits", "max_num_prims_per_leaf"):
        assert st[k] == ref_stats[k], k
    assert st["num_primitives"] == 6000


class _Boxes:
    """A host collection in the reference's shape (GEOM_TEMPLATE.h:60-96)."""

    def __init__(self, lo, hi, ctr):
        self.lo, self.hi, self.ctr = lo, hi, ctr

    def size(self):
        return len(self.lo)

    def get_primitive_bounding_box(self, i):
        return self.lo[i], self.hi[i]

    def get_primitive_center(self, i):
        return self.ctr[i]


def test_collection_object_equals_arrays():
    lo, hi, ctr = _boxes(500, 12)
    a = BVH(_Boxes(lo, hi, ctr))
    b = BVH(box_min=lo, box_max=hi, centers=ctr)
    _same_tree(a.export_tree()[0], b.export_tree()[0])
    assert np.array_equal(a.export_tree()[1], b.export_tree()[1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sphere_boxes_give_the_scene_tree(dtype):
    """A collection with the sphere boxes (center -+ radius, sphere.h:88-99) gets the tree the built-in sphere type
    gets, so the plugin's traversal of it can be compared with the built-in sphere path hit for hit."""
    c, r = W.sphere_field(3000, seed=13)
    c, r = c.astype(dtype), r.astype(dtype)
    lo, hi = c - r[:, None], c + r[:, None]
    b = BVH(box_min=lo, box_max=hi, centers=c, dtype=dtype)
    s = Scene(dtype=dtype)
    s.add_spheres(c, r)
    s.commit()
    sn, so = s.export_tree(1)
    bn, bo = b.export_tree()
    _same_tree(bn, sn)
    assert np.array_equal(bo, so)


def test_empty_and_invalid():
    z = np.zeros((0, 3))
    b = BVH(box_min=z, box_max=z, centers=z)
    nodes, order = b.export_tree()
    assert len(nodes) == 1 and nodes["leaf"][0] == 1 and nodes["prim_count"][0] == 0 and len(order) == 0
    with pytest.raises(ValueError):
        BVH(box_min=np.zeros((3, 3)), box_max=np.zeros((2, 3)), centers=np.zeros((3, 3)))
    with pytest.raises(BlazertError):
        BVH(box_min=np.zeros((3, 3)), box_max=np.zeros((3, 3)), centers=np.zeros((3, 3)), bin_size=1)
    with pytest.raises(BlazertError):
        BVH(box_min=np.zeros((3, 3)), box_max=np.zeros((3, 3)), centers=np.zeros((3, 3)), min_leaf_primitives=0)


def test_device_view_needs_a_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    lo, hi, ctr = _boxes(10, 14)
    with pytest.raises(BlazertError) as e:
        BVH(box_min=lo, box_max=hi, centers=ctr).device_view()
    assert e.value.code == -4


def test_plugin_library_and_template_builder(oracle_mod):
    """The example plugin (examples/user_geometry/user_geometry.hip) loads and exports its entry points; its
    BVH<double> built from a host collection object through the header's template constructor is the oracle's
    tree."""
    assert os.path.exists(PLUGIN), "build the plugin first (__graft_entry__.build())"
    L = C.CDLL(PLUGIN)
    for name in ("ug_spheres_build", "ug_spheres_intersect", "ug_spheres_intersect_f32", "ug_boxes_intersect"):
        assert hasattr(L, name), name
    from blazert_amd import capi
    c, r = W.sphere_field(2000, seed=15)
    c, r = np.ascontiguousarray(c), np.ascontiguousarray(r)
    dp = C.POINTER(C.c_double)
    L.ug_spheres_build.argtypes = [dp, dp, C.c_uint, C.POINTER(capi.BuildOptions), C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    opts = capi.BuildOptions(4, 28, 64)
    assert L.ug_spheres_build(c.ctypes.data_as(dp), r.ctypes.data_as(dp), len(r), C.byref(opts), C.byref(h)) == 0
    try:
        n = C.c_uint32(0)
        capi.check(capi.lib().brt_bvh_export_tree(h, None, 0, None, C.byref(n)))
        from blazert_amd.collection import TREE_DTYPE
        nodes = np.zeros(n.value, TREE_DTYPE)
        order = np.zeros(len(r), np.uint32)
        capi.check(capi.lib().brt_bvh_export_tree(h, nodes.ctypes.data_as(C.c_void_p), n.value,
                                                  order.ctypes.data_as(C.POINTER(C.c_uint32)), C.byref(n)))
    finally:
        capi.lib().brt_bvh_destroy(h)
    ref_nodes, ref_order, _ = oracle_mod.collection_tree(c - r[:, None], c + r[:, None], c)
    _same_tree(nodes, ref_nodes)
    assert np.array_equal(order, ref_order)
