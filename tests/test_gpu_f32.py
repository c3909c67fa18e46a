"""blazert::Scene<float> on the GPU (brt_scene_create_f32 / brt_intersect_f32 / brt_intersect1_f32) against the FP32
oracle (oracle/liboracle_f32.so, pinned by the reference's float-templated known answers).

Every field (t, u, v, prim_id, geom_id, normal) must be bit-identical to the oracle's T = float path, on every ray
whose reference traversal stays within its 28-entry stack; the build must give the oracle's tree.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401
    return torch.device("cuda:0")


def _pair(oracle, fill, **opts):
    from blazert_amd import Scene
    g = Scene(dtype=np.float32, **opts)
    fill(g)
    g.commit()
    o = oracle.SceneF32(*(opts.get(k, d) for k, d in (("min_leaf_primitives", 4), ("max_tree_depth", 28),
                                                      ("bin_size", 64))))
    fill(o)
    o.commit()
    return g, o


def _gpu(scene, org, dirs, dev, **kw):
    O = torch.from_numpy(np.ascontiguousarray(org, dtype=np.float32)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(dirs, dtype=np.float32)).to(dev)
    h = scene.intersect(O, D, normals=True, **kw)
    torch.cuda.synchronize()
    out = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
    out["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
    out["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
    out["normal"] = h["normal"].cpu().numpy().T
    return out


def _assert_same(g, r, mask=None):
    m = np.ones(len(g["t"]), bool) if mask is None else mask
    for k in ("t", "u", "v", "normal"):
        a, b = np.asarray(g[k])[m], np.asarray(r[k])[m]
        same = (a.view(np.uint32) == b.view(np.uint32)) | (np.isnan(a) & np.isnan(b))
        assert same.all(), f"{k}: {int((~same).sum())} mismatches"
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(np.asarray(g[k])[m], np.asarray(r[k])[m]), k


def _fill_case(case):
    def fill(s):
        for g in case.get("add_order") or [case["geometry"]]:
            if g == "mesh":
                s.add_mesh(np.array(case["mesh"]["vertices"], float), np.array(case["mesh"]["faces"], np.uint32))
            elif g == "spheres":
                s.add_spheres(np.array(case["spheres"]["centers"], float), case["spheres"]["radii"])
            elif g == "planes":
                d = case["planes"]
                s.add_planes(np.array(d["centers"], float), d["dxs"], d["dys"], np.array(d["rotations"], float))
            else:
                d = case["cylinders"]
                s.add_cylinders(np.array(d["centers"], float), d["a"], d["b"], d["heights"],
                                np.array(d["rotations"], float))
    return fill


def test_known_answers_f32(dev, oracle_mod):
    doc = json.load(open(GOLDEN))
    n = 0
    for case in doc["cases"]:
        if case["kind"] not in ("primitive", "scene", "traverse", "traverse_distance"):
            continue
        g, o = _pair(oracle_mod, _fill_case(case))
        org, d = np.array([case["org"]], float), np.array([case["dir"]], float)
        _assert_same(_gpu(g, org, d, dev), o.intersect_batch(org, d))
        n += 1
    assert n > 60


@pytest.mark.parametrize("level", [2, 6])
def test_icosphere_f32(dev, oracle_mod, level):
    V, F = W.icosphere(level)
    g, o = _pair(oracle_mod, lambda s: s.add_mesh(V, F))
    nodes_g, order_g = g.export_tree(0)
    nodes_o, order_o, _ = o.tree(0)
    assert np.array_equal(order_g, order_o)
    assert np.array_equal(nodes_g["min"].astype(np.float32), nodes_o["min"])
    assert np.array_equal(nodes_g["children"], nodes_o["children"])
    org, dirs = W.incoherent_rays(1 << 15, 5, 2.0)
    print(f"Scene<float>: icosphere level {level}, {len(F)} triangles", flush=True)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    _assert_same(_gpu(g, org, dirs, dev), ref, ok)
    corg, cdirs = W.camera_rays(160, 120, (0.0, 0.0, 4.0))
    _assert_same(_gpu(g, corg, cdirs, dev, schedule=1), o.intersect_batch(corg, cdirs))


def test_mixed_scene_f32(dev, oracle_mod):
    m = W.mixed_scene(20_000, 2_000, 500, 50, seed=17)

    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    g, o = _pair(oracle_mod, fill)
    org, dirs = W.mixed_rays(1 << 15, seed=18)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    print(f"Scene<float>: mixed scene, {int(ok.sum())} rays compared", flush=True)
    _assert_same(_gpu(g, org, dirs, dev), ref, ok)
    # any-hit and tmin / tmax windows
    ra = o.intersect_batch(org, dirs, any_hit=True, trace=True)
    _assert_same(_gpu(g, org, dirs, dev, any_hit=True), ra, ra["trace"]["overflow"] == 0)
    tmin = np.where(W.Rng(19).uniform(len(org)) < 0.3, 2.0, 0.0).astype(np.float32)
    tmax = np.where(W.Rng(20).uniform(len(org)) < 0.3, 8.0, np.finfo(np.float32).max).astype(np.float32)
    rt = o.intersect_batch(org, dirs, tmin=tmin, tmax=tmax, trace=True)
    g2 = _gpu(g, org, dirs, dev, tmin=torch.from_numpy(tmin).to(dev), tmax=torch.from_numpy(tmax).to(dev))
    _assert_same(g2, rt, rt["trace"]["overflow"] == 0)


def test_intersect1_f32(dev, oracle_mod):
    from blazert_amd import Ray, RayHit, intersect1
    c, r = W.sphere_field(64, seed=3)
    g, o = _pair(oracle_mod, lambda s: s.add_spheres(c, r))
    rays = W.incoherent_rays(64, 21, 10.0)
    for k in range(64):
        rh = RayHit()
        hit = intersect1(g, Ray(tuple(rays[0][k]), tuple(rays[1][k])), rh)
        ok, h, _ = o.intersect1(oracle_mod.make_ray_f32(rays[0][k], rays[1][k]))
        assert hit == ok
        assert np.float32(rh.hit_distance).view(np.uint32) == np.float32(h.t).view(np.uint32)
        assert rh.prim_id == h.prim_id and rh.geom_id == h.geom_id
