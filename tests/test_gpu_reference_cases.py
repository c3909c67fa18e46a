"""Every intersection known answer of the reference's primitive test suite, on the GPU through the C ABI.

tests/golden/reference_cases.json holds each assert of blazert's test/test_primitives/test_{trimesh,sphere,plane,
cylinder}.cpp (extracted by tests/golden/make_reference_cases.py). The rays of each distinct collection are traced
in one launch:

* FP64 path: every field bit-identical to the oracle, and the reference's expected hit flag and distance (doctest
  Approx) for the traversal asserts;
* split precision (FP32 traversal, FP64 hit): no disagreement with the FP64 oracle outside the classes the split
  path's error bound allows (ties, cylinders), and every agreeing ray bit-identical;
* Scene<float>: every field bit-identical to the FP32 oracle.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")
DOC = json.load(open(GOLDEN))
EPS = DOC["approx_epsilon"]
INVALID = 0xFFFFFFFF
RAY_ASSERTS = ("intersect_primitive", "traverse", "traverse_distance", "traverse_collection")


def groups():
    g = {}
    for c in DOC["cases"]:
        if c["assert"] in RAY_ASSERTS:
            g.setdefault(json.dumps(c["scene"], sort_keys=True), []).append(c)
    return [pytest.param(v, id=f"{v[0]['ref'].split('/')[-1]}+{len(v)}") for v in g.values()]


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401  (raises if libblazert_amd.so is missing)
    return torch.device("cuda:0")


def fill(s, sc):
    g = sc["geometry"]
    rot = lambda m: np.array(m, float).reshape(-1, 9)
    if g == "mesh":
        return s.add_mesh(np.array(sc["vertices"], float).reshape(-1, 3), np.array(sc["faces"], np.uint32).reshape(-1, 3))
    if g == "spheres":
        return s.add_spheres(np.array(sc["centers"], float).reshape(-1, 3), np.array(sc["radii"], float))
    if g == "planes":
        return s.add_planes(np.array(sc["centers"], float).reshape(-1, 3), sc["dxs"], sc["dys"], rot(sc["rotations"]))
    return s.add_cylinders(np.array(sc["centers"], float).reshape(-1, 3), sc["a"], sc["b"], sc["heights"],
                           rot(sc["rotations"]))


def trace(scene, org, dirs, dev, dtype=np.float64, **kw):
    O = torch.from_numpy(np.ascontiguousarray(org, dtype=dtype)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(dirs, dtype=dtype)).to(dev)
    h = scene.intersect(O, D, normals=True, **kw)
    torch.cuda.synchronize()
    out = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
    out["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
    out["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
    out["normal"] = h["normal"].cpu().numpy().T
    return out


def same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype.kind != "f":
        return a == b
    u = np.uint64 if a.dtype == np.float64 else np.uint32
    eq = (a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))
    return eq.all(axis=1) if eq.ndim == 2 else eq


def rays_of(cases):
    return np.array([c["org"] for c in cases], float), np.array([c["dir"] for c in cases], float)


@pytest.mark.parametrize("cases", groups())
def test_reference_cases_fp64(dev, oracle_mod, cases):
    from blazert_amd import Scene
    sc = cases[0]["scene"]
    p, o = Scene(), oracle_mod.Scene()
    fill(p, sc)
    fill(o, sc)
    p.commit()
    o.commit()
    org, dirs = rays_of(cases)
    g = trace(p, org, dirs, dev)
    r = o.intersect_batch(org, dirs)
    for k in ("t", "u", "v", "prim_id", "geom_id", "normal"):
        assert same_bits(g[k], r[k]).all(), (k, [c["ref"] for c in cases])
    for i, c in enumerate(cases):
        if c["assert"] == "intersect_primitive":
            continue  # primitive 0 tested alone: the oracle pins it (tests/test_oracle_reference_cases.py)
        assert (g["geom_id"][i] != INVALID) == c["hit"], c["ref"]
        if c["hit"]:
            assert abs(g["t"][i] - c["t"]) < EPS * (1 + max(abs(g["t"][i]), abs(c["t"]))), c["ref"]


@pytest.mark.parametrize("cases", groups())
def test_reference_cases_split_precision(dev, oracle_mod, cases):
    from blazert_amd import PRECISION_FP32_TRAV_FP64_HIT, Scene
    sc = cases[0]["scene"]
    p, o = Scene(), oracle_mod.Scene()
    gid = fill(p, sc)
    fill(o, sc)
    p.commit()
    o.commit()
    org, dirs = rays_of(cases)
    g = trace(p, org, dirs, dev, precision=PRECISION_FP32_TRAV_FP64_HIT)
    r = o.intersect_batch(org, dirs)
    agree = (g["prim_id"] == r["prim_id"]) & (g["geom_id"] == r["geom_id"])
    idx = np.nonzero(~agree)[0]
    cls = oracle_mod.classify_split_disagreements(o, org, dirs, idx, g["t"][idx], g["geom_id"][idx],
                                                  cyl_geom=gid if sc["geometry"] == "cylinders" else None)
    assert cls["unexplained"] == 0, [cases[i]["ref"] for i in cls["unexplained_idx"]]
    for k in ("t", "u", "v", "normal"):
        assert same_bits(g[k][agree], r[k][agree]).all(), k


@pytest.mark.parametrize("cases", groups())
def test_reference_cases_f32(dev, oracle_mod, cases):
    from blazert_amd import Scene
    sc = cases[0]["scene"]
    p, o = Scene(dtype=np.float32), oracle_mod.SceneF32()
    fill(p, sc)
    fill(o, sc)
    p.commit()
    o.commit()
    org, dirs = rays_of(cases)
    g = trace(p, org, dirs, dev, dtype=np.float32)
    r = o.intersect_batch(org.astype(np.float32), dirs.astype(np.float32))
    for k in ("t", "u", "v", "prim_id", "geom_id", "normal"):
        assert same_bits(g[k], r[k]).all(), (k, [c["ref"] for c in cases])
