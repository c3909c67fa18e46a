"""Host-side checks of the extensions (no GPU): argument validation of traverse, mirror tracing and the autotuner,
the autotuner's cache, and the synthetic workload generators the benchmark uses."""
import ctypes as C
import json

import numpy as np
import pytest

from blazert_amd import Scene, capi
from blazert_amd import workloads as W


def _scene_with(kind):
    s = Scene()
    if kind == "spheres":
        s.add_spheres(np.zeros((1, 3)), [1.0])
    else:
        v, f = W.icosphere(1)
        s.add_mesh(v, f)
    s.commit()
    return s


def test_stack_ring_validated_on_host():
    s = _scene_with("spheres")
    rays = capi.Rays(1, 1, 1, 1, 1, 1, None, None)  # never dereferenced: validation fails first
    hits = capi.Hits(1, 1, 1, 1, 1, None, None, None)
    for ring in (1, 2, 3, 12, 32):
        L = capi.Launch(0, 0, 0, 0, 0, 0, ring)
        rc = capi.lib().brt_intersect(s.handle, C.byref(rays), C.byref(hits), 10, C.byref(L), None)
        assert rc == capi.ERR_INVALID_ARGUMENT, ring


def test_traverse_validation():
    s = _scene_with("spheres")
    rays = capi.Rays(1, 1, 1, 1, 1, 1, None, None)
    hits = capi.Hits(1, 1, 1, 1, 1, None, None, None)
    L = capi.Launch(0, 0, 0, 0, 0, 0)
    assert capi.lib().brt_traverse(s.handle, 0, C.byref(rays), C.byref(hits), 10, C.byref(L), None) == \
        capi.ERR_INVALID_ARGUMENT  # no triangles in this scene
    assert capi.lib().brt_traverse(s.handle, 9, C.byref(rays), C.byref(hits), 10, C.byref(L), None) == \
        capi.ERR_INVALID_ARGUMENT
    L.flags = capi.FLAG_HITS_INOUT
    assert capi.lib().brt_traverse(s.handle, 1, C.byref(rays), C.byref(hits), 10, C.byref(L), None) == \
        capi.ERR_INVALID_ARGUMENT
    L.flags, L.sort = 0, capi.SORT_MORTON
    assert capi.lib().brt_traverse(s.handle, 1, C.byref(rays), C.byref(hits), 10, C.byref(L), None) == \
        capi.ERR_UNSUPPORTED


def test_mirror_validation():
    rays = capi.Rays(1, 1, 1, 1, 1, 1, None, None)
    out = capi.MirrorOut(*([1] * 10))
    sph = _scene_with("spheres")
    tri = _scene_with("mesh")
    lib = capi.lib()
    assert lib.brt_mirror_trace(sph.handle, C.byref(rays), 4, 3, 1e-9, None, C.byref(out), None) == capi.ERR_UNSUPPORTED
    assert lib.brt_mirror_trace(tri.handle, C.byref(rays), 4, 0, 1e-9, None, C.byref(out), None) == \
        capi.ERR_INVALID_ARGUMENT
    assert lib.brt_mirror_trace(tri.handle, C.byref(rays), 4, 9, 1e-9, None, C.byref(out), None) == \
        capi.ERR_INVALID_ARGUMENT
    assert lib.brt_mirror_trace(tri.handle, C.byref(rays), 4, 3, -1.0, None, C.byref(out), None) == \
        capi.ERR_INVALID_ARGUMENT
    L = capi.Launch(capi.PRECISION_FP32_TRAV_FP64_HIT, 0, 0, 0, 0, capi.SORT_MORTON)  # mirror bounces keep ray order
    assert lib.brt_mirror_trace(tri.handle, C.byref(rays), 4, 3, 1e-9, C.byref(L), C.byref(out), None) == \
        capi.ERR_UNSUPPORTED
    g = capi.MirrorGrads(None, 1, None, None, None, None, None, None, None, None, None)  # gp x without y, z
    assert lib.brt_mirror_adjoint(tri.handle, C.byref(rays), 4, 3, 1, C.byref(g), None) == capi.ERR_INVALID_ARGUMENT


def test_autotune_validation():
    s = _scene_with("spheres")
    rays = capi.Rays(1, 1, 1, 1, 1, 1, None, None)
    hits = capi.Hits(1, 1, 1, 1, 1, None, None, None)
    r = capi.TuneResult()
    lib = capi.lib()
    assert lib.brt_autotune_launch(s.handle, C.byref(rays), C.byref(hits), 0, 0, 0, 2, C.byref(r), None) == \
        capi.ERR_INVALID_ARGUMENT
    assert lib.brt_autotune_launch(s.handle, C.byref(rays), C.byref(hits), 8, 0, capi.FLAG_HITS_INOUT, 2, C.byref(r),
                                   None) == capi.ERR_INVALID_ARGUMENT
    sizes = (C.c_uint32 * 2)(2, 4)
    assert lib.brt_scene_tune_leaf(s.handle, C.byref(rays), C.byref(hits), 8, 0, sizes, 0, 2, C.byref(r), None) == \
        capi.ERR_INVALID_ARGUMENT


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    from blazert_amd import autotune
    monkeypatch.setenv("BLAZERT_AMD_TUNE_CACHE", str(tmp_path / "sub" / "tune.json"))
    assert autotune.cache_path() == str(tmp_path / "sub" / "tune.json")
    assert autotune._load() == {}
    autotune._store("k1", {"block_size": 128})
    autotune._store("k2", {"block_size": 64})
    data = json.load(open(tmp_path / "sub" / "tune.json"))
    assert data == {"k1": {"block_size": 128}, "k2": {"block_size": 64}}
    assert [autotune._bucket(n) for n in (1, 3, 1000, 1 << 20, (1 << 20) + 5)] == [1, 4, 1024, 1 << 20, 1 << 20]
    s = _scene_with("mesh")
    assert autotune.scene_signature(s).startswith("0:80:")  # 80 triangles at level 1


def test_torus_is_closed_and_sized():
    for target in (1000, 20000, 123457):
        v, f = W.torus(target, seed=3)
        assert abs(len(f) - target) / target < 0.1
        edges = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, counts = np.unique(edges, axis=0, return_counts=True)
        assert (counts == 2).all()  # every edge shared by exactly two triangles


def test_uneven_batch_and_mirror_workloads():
    b = W.uneven_batch(count=6, seed=3, tri_range=(1000, 20000), ray_range=(1000, 5000))
    assert len(b) == 6
    for it in b:
        assert 1000 * 0.3 <= len(it["faces"]) <= 20000 * 4  # icosphere sizes snap to powers of four
        assert 1000 <= len(it["origins"]) <= 5000
        assert np.allclose(np.linalg.norm(it["directions"], axis=1), 1.0)
    v, f = W.mirror_box_scene(level=2, half=3.0)
    assert len(f) == 320 + 12 and len(v) == 162 + 8
    assert np.abs(v[-8:]).max() == 3.0
    w = W.mirror_weights(100, 3)
    assert w.shape == (3, 100) and (w >= 0).all() and (w < 1).all()


def test_scene_f32_type_guards():
    """Scene<float> takes the _f32 adds only, and refuses the Scene<double>-only entry points before any GPU work."""
    import ctypes as C

    from blazert_amd import Scene, capi
    from blazert_amd.capi import lib
    s = Scene(dtype=np.float32)
    assert s.add_spheres(np.zeros((2, 3)), [1.0, 2.0]) == 0
    d = np.zeros(3)
    rc = lib().brt_scene_add_mesh(s.handle, d.ctypes.data_as(C.POINTER(C.c_double)), 1,
                                  np.zeros(3, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32)), 0, None)
    assert rc == capi.ERR_INVALID_ARGUMENT
    s.commit()
    st = s.build_stats(1)
    assert st["num_primitives"] == 2 and st["built_on_device"] == 0
    nodes, order = s.export_tree(1)
    assert len(nodes) >= 1 and sorted(order) == [0, 1]
    rays = capi.Rays()
    hits = capi.Hits()
    assert lib().brt_intersect(s.handle, C.byref(rays), C.byref(hits), 1, None, None) == capi.ERR_UNSUPPORTED
    assert lib().brt_scene_commit_device(s.handle, None) == capi.ERR_UNSUPPORTED
    sd = Scene()
    f = np.zeros(3, np.float32)
    rc = lib().brt_scene_add_spheres_f32(sd.handle, f.ctypes.data_as(C.POINTER(C.c_float)),
                                         f.ctypes.data_as(C.POINTER(C.c_float)), 1, None)
    assert rc == capi.ERR_INVALID_ARGUMENT
