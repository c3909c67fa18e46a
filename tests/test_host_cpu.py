"""Host-side checks that need no GPU: the C ABI library loads and exports every declared entry point, the scene
API keeps blazeRT's add/commit semantics, and the product builder emits exactly the oracle's tree (topology,
boxes, split axes, leaf primitive order and statistics)."""
import os
import re

import numpy as np
import pytest

from blazert_amd import capi, workloads as W
from blazert_amd.scene import Scene

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "blazert_amd.h")).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(brt_[a-z_0-9]+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    L = capi.lib()
    names = declared_symbols()
    assert len(names) >= 14, names
    for n in names:
        assert hasattr(L, n), n
        assert n in capi.SIGNATURES, n
    assert set(capi.SIGNATURES) == set(names)


def test_version_and_device_count_without_gpu():
    assert b"gfx950" in capi.lib().brt_version()
    assert capi.lib().brt_device_count() >= 0


def test_add_semantics_match_reference():
    s = Scene()
    assert s.add_spheres(np.zeros((2, 3)), [1.0, 2.0]) == 0
    assert s.add_spheres(np.zeros((1, 3)), [1.0]) == capi.INVALID_ID  # second add of a type
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    assert s.add_mesh(v, [[0, 1, 2]]) == 1  # geom ids follow add order
    s.commit()
    assert s.add_planes(np.zeros((1, 3)), [1.0], [1.0], np.eye(3).reshape(1, 9)) == capi.INVALID_ID


def test_bad_face_index_rejected():
    s = Scene()
    with pytest.raises(capi.BlazertError):
        s.add_mesh(np.zeros((3, 3)), [[0, 1, 3]])


def _oracle_scene(oracle, build):
    o = oracle.Scene()
    build(o)
    o.commit()
    return o


def _compare_trees(prod, orc, t):
    pn, po = prod.export_tree(t)
    on, oo, ost = orc.tree(t)
    assert len(pn) == len(on)
    assert np.array_equal(pn["min"].view(np.uint64), on["min"].view(np.uint64))
    assert np.array_equal(pn["max"].view(np.uint64), on["max"].view(np.uint64))
    assert np.array_equal(pn["leaf"], on["leaf"])
    internal = on["leaf"] == 0
    assert np.array_equal(pn["axis"][internal], on["axis"][internal])
    assert np.array_equal(pn["children"][internal], on["children"][internal])
    leaf = ~internal
    assert np.array_equal(pn["prim_begin"][leaf], on["prim_begin"][leaf])
    assert np.array_equal(pn["prim_count"][leaf], on["prim_count"][leaf])
    assert np.array_equal(po, oo)
    ps = prod.build_stats(t)
    for k in ("max_tree_depth", "leaf_nodes", "max_num_prims_per_leaf", "min_num_prims_per_leaf", "branch_nodes",
              "bad_splits"):
        assert ps[k] == ost[k], k
    return pn, po


def _mesh_random(n, seed):
    r = W.Rng(seed)
    c = r.uniform(3 * n, -5, 5).reshape(n, 3)
    v = np.concatenate([c + r.uniform(3 * n, -0.3, 0.3).reshape(n, 3) for _ in range(3)])
    f = np.stack([np.arange(n), np.arange(n) + n, np.arange(n) + 2 * n], 1).astype(np.uint32)
    return v, f


@pytest.mark.parametrize("n,seed", [(1, 1), (4, 2), (5, 3), (100, 4), (5000, 5)])
def test_builder_matches_oracle_random_triangles(oracle_mod, n, seed):
    v, f = _mesh_random(n, seed)
    p = Scene()
    p.add_mesh(v, f)
    p.commit()
    o = _oracle_scene(oracle_mod, lambda s: s.add_mesh(v, f))
    _compare_trees(p, o, 0)


def test_builder_matches_oracle_icosphere_parallel(oracle_mod):
    v, f = W.icosphere(6)  # 81,920 triangles: deep enough for the threaded subtree build
    p = Scene()
    p.add_mesh(v, f)
    p.commit()
    o = _oracle_scene(oracle_mod, lambda s: s.add_mesh(v, f))
    nodes, _ = _compare_trees(p, o, 0)
    assert len(nodes) > 30000


def test_builder_matches_oracle_options(oracle_mod):
    v, f = W.icosphere(4)
    p = Scene(min_leaf_primitives=2, max_tree_depth=7, bin_size=16)
    p.add_mesh(v, f)
    p.commit()
    o = oracle_mod.Scene(2, 7, 16)
    o.add_mesh(v, f)
    o.commit()
    _compare_trees(p, o, 0)
    assert p.build_stats(0)["max_tree_depth"] == 7


def test_builder_matches_oracle_mixed(oracle_mod):
    m = W.mixed_scene(n_tri_target=20000, n_sph=3000, n_cyl=1000, n_pla=50, seed=5)

    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_planes(*m["planes"])
        s.add_cylinders(*m["cylinders"])

    p = Scene()
    fill(p)
    p.commit()
    o = _oracle_scene(oracle_mod, fill)
    for t in range(4):
        _compare_trees(p, o, t)


def test_builder_degenerate_collections(oracle_mod):
    # all centers identical -> every split is bad; median fallback must match
    c = np.zeros((37, 3))
    r = np.linspace(0.1, 1.0, 37)
    p = Scene()
    p.add_spheres(c, r)
    p.commit()
    o = _oracle_scene(oracle_mod, lambda s: s.add_spheres(c, r))
    _compare_trees(p, o, 1)
    assert p.build_stats(1)["bad_splits"] > 0


def test_tree_invariants():
    v, f = W.icosphere(5)
    p = Scene()
    p.add_mesh(v, f)
    p.commit()
    nodes, order = p.export_tree(0)
    assert np.array_equal(np.sort(order), np.arange(len(f)))  # every primitive in exactly one leaf
    internal = np.nonzero(nodes["leaf"] == 0)[0]
    assert len(nodes) == 2 * (len(nodes) - len(internal)) - 1
    for i in internal:
        l, r = nodes["children"][i]
        assert l == i + 1  # pre-order: left child follows its parent
        lo = np.minimum(nodes["min"][l], nodes["min"][r])
        hi = np.maximum(nodes["max"][l], nodes["max"][r])
        assert np.array_equal(lo, nodes["min"][i]) and np.array_equal(hi, nodes["max"][i])
    leaves = nodes[nodes["leaf"] == 1]
    assert (leaves["prim_count"] <= 4).all() or p.build_stats(0)["max_tree_depth"] >= 28


def test_launch_options_validated_on_host():
    """Invalid launch options are rejected before any device work (block sizes above the kernels' bound)."""
    import ctypes as C
    s = Scene()
    s.add_spheres(np.zeros((1, 3)), [1.0])
    s.commit()
    rays = capi.Rays(1, 1, 1, 1, 1, 1, None, None)  # never dereferenced: validation fails first
    hits = capi.Hits(1, 1, 1, 1, 1, None, None, None)
    for L in (capi.Launch(0, 0, 512, 0, 0, 0), capi.Launch(0, 0, 96, 0, 0, 0), capi.Launch(7, 0, 0, 0, 0, 0),
              capi.Launch(0, 0, 0, 5, 0, 0), capi.Launch(0, 0, 0, 0, 65, 0), capi.Launch(0, 0, 0, 0, 0, 9),
              capi.Launch(0, 0, 0, 0, 0, 0, 0, 100), capi.Launch(0, 4, 0, 0, 0, 0)):
        rc = capi.lib().brt_intersect(s.handle, C.byref(rays), C.byref(hits), 10, C.byref(L), None)
        assert rc == capi.ERR_INVALID_ARGUMENT, (L.block_size, rc)


def test_microbench_library_exports():
    """The bench's ceiling microbenchmarks (blazert_amd/csrc/microbench.hip) build into their own library, apart from
    the ray-query library, and export their three entry points."""
    import ctypes as C
    from blazert_amd import build as B
    if not os.path.exists(B.MICROBENCH_LIB):
        B.build_microbench()
    lib = C.CDLL(B.MICROBENCH_LIB)
    for name in ("brt_mb_line_chase", "brt_mb_line_gather", "brt_mb_fp64_atomics"):
        assert hasattr(lib, name), name
    assert not hasattr(capi.lib(), "brt_mb_line_chase")
