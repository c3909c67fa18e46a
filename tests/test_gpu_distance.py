"""GPU distance queries (brt_distance_to_surface, brt_closest_surface) against the CPU oracle, through the C ABI.

Per-pair distances are bit-exact for every geometry type (the oracle restates the reference's arithmetic and is
pinned by its cylinder known answers, geometric constructions and an independent closest-point construction). The
closest-surface query must equal the oracle's brute-force minimum bit for bit, including which primitive wins ties.
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
TYPES = {0: "mesh", 1: "spheres", 2: "planes", 3: "cylinders"}


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401
    return torch.device("cuda:0")


def _fill(m):
    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    return fill


@pytest.fixture(scope="module")
def mixed(oracle_mod):
    from blazert_amd import Scene
    m = W.mixed_scene(20_000, 2_000, 500, 50, seed=21)
    g = Scene()
    _fill(m)(g)
    g.commit()
    o = oracle_mod.Scene()
    _fill(m)(o)
    o.commit()
    return m, g, o


def _same(a, b):
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) or \
        bool(((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))).all())


@pytest.mark.parametrize("type_", [0, 1, 2, 3])
def test_pairs_bit_exact(dev, mixed, type_):
    m, g, o = mixed
    count = len(m[TYPES[type_]][1])
    n = 1 << 14
    pts = W.incoherent_rays(n, 30 + type_, 22.0)[0]
    ids = (W.Rng(40 + type_).uniform(n) * count).astype(np.uint32)
    ids[:16] = count + np.arange(16)  # out of range -> NaN
    print(f"distance pairs: type {type_}, {n} points", flush=True)
    got = g.distance_to_surface(type_, torch.from_numpy(pts).to(dev), torch.from_numpy(ids.view(np.int32)).to(dev))
    torch.cuda.synchronize()
    ref = o.distance_to_surface(type_, pts, ids)
    got = got.cpu().numpy()
    assert np.isnan(got[:16]).all()
    assert _same(got, ref)


def test_cylinder_known_answers_on_gpu(dev, oracle_mod):
    from blazert_amd import Scene
    doc = json.load(open(os.path.join(GOLDEN, "cylinder_distance_cases.json")))
    eps = doc["approx_epsilon"]
    groups = {}
    for c in doc["cases"]:
        key = (tuple(c["center"]), c["a"], c["b"], c["height"], tuple(c["rotation"]))
        groups.setdefault(key, []).append(c)
    for (ctr, a, b, h, R), cases in groups.items():
        g = Scene()
        g.add_cylinders(np.array([ctr]), [a], [b], [h], np.array([R]))
        g.commit()
        pts = np.array([c["point"] for c in cases])
        got = g.distance_to_surface(3, torch.from_numpy(pts).to(dev), torch.zeros(len(cases), dtype=torch.int32, device=dev))
        got = got.cpu().numpy()
        for c, d in zip(cases, got):
            exp = c["distance"]
            assert abs(d - exp) <= eps * (1 + max(abs(d), abs(exp))), (c["line"], d, exp)
            ref = oracle_mod.cylinder_distance(c["center"], c["a"], c["b"], c["height"], c["rotation"], c["point"])
            assert np.float64(d).view(np.uint64) == np.float64(ref).view(np.uint64), (c["line"], d, ref)


def _closest_check(g, o, pts, dev):
    out = g.closest_surface(torch.from_numpy(pts).to(dev))
    torch.cuda.synchronize()
    ref = o.closest_surface(pts)
    d = out["distance"].cpu().numpy()
    assert _same(d, ref["distance"])
    assert np.array_equal(out["prim_id"].cpu().numpy().view(np.uint32), ref["prim_id"])
    assert np.array_equal(out["geom_id"].cpu().numpy().view(np.uint32), ref["geom_id"])
    return d


def test_closest_surface_mixed(dev, mixed):
    m, g, o = mixed
    pts = W.incoherent_rays(2048, 50, 22.0)[0]
    print("closest surface: mixed scene, 2048 points", flush=True)
    _closest_check(g, o, pts, dev)


@pytest.mark.parametrize("device_commit", [False, True])
def test_closest_surface_icosphere(dev, oracle_mod, device_commit):
    from blazert_amd import Scene
    V, F = W.icosphere(5)
    g = Scene()
    g.add_mesh(V, F)
    g.commit(device=device_commit)
    o = oracle_mod.Scene()
    o.add_mesh(V, F)
    o.commit()
    r = W.Rng(51)
    u = r.unit_vectors(1500)
    rad = np.concatenate([r.uniform(500, 0.0, 0.9), r.uniform(500, 0.97, 1.03), r.uniform(500, 1.5, 6.0)])
    pts = u * rad[:, None]
    print(f"closest surface: icosphere level 5 ({len(F)} triangles), device commit {device_commit}", flush=True)
    d = _closest_check(g, o, pts, dev)
    assert np.isfinite(d).all()


def test_closest_surface_inverted_sphere_boxes(dev, oracle_mod):
    """Spheres with negative radii have inverted boxes, which box pruning cannot use: such a tree is scanned in full,
    and the result must still be the brute-force minimum."""
    from blazert_amd import Scene
    c, r = W.sphere_field(3000, seed=46)
    r = np.where(W.Rng(47).uniform(len(r)) < 0.3, -r, r)
    V, F = W.icosphere(3)

    def fill(s):
        s.add_mesh(V * 4.0, F)
        s.add_spheres(c, r)
    g = Scene()
    fill(g)
    g.commit()
    o = oracle_mod.Scene()
    fill(o)
    o.commit()
    pts = W.incoherent_rays(2048, 48, 12.0)[0]
    print("closest surface: inverted sphere boxes, 2048 points", flush=True)
    d = _closest_check(g, o, pts, dev)
    assert np.isfinite(d).all()


def test_closest_surface_ties_and_empty(dev, oracle_mod):
    from blazert_amd import Scene

    def fill(s):
        s.add_spheres(np.array([[0.0, 0, 0], [0.0, 0, 0], [5.0, 0, 0]]), [1.0, 1.0, 1.0])
        s.add_mesh(np.array([[-3.0, -1, 2], [3.0, -1, 2], [0.0, 3, 2]]), np.array([[0, 1, 2]], np.uint32))
    g = Scene()
    fill(g)
    g.commit()
    o = oracle_mod.Scene()
    fill(o)
    o.commit()
    pts = np.array([[0.0, 0.0, 1.5], [0.0, 0.0, -1.5], [5.0, 0.0, 0.0], [100.0, 0, 0]])
    _closest_check(g, o, pts, dev)
    e = Scene()
    e.add_spheres(np.zeros((0, 3)), np.zeros(0))
    e.commit()
    out = e.closest_surface(torch.from_numpy(pts).to(dev))
    assert np.isinf(out["distance"].cpu().numpy()).all()
    assert (out["prim_id"].cpu().numpy() == -1).all()
