"""GPU parity: the HIP path (through the C ABI) against the CPU oracle on identical seeded inputs.

FP64 path: bit-exact on every field (t, u, v, prim_id, geom_id, normal) for every ray except those where the
reference's fixed 28-slot stack would have overflowed (undefined behaviour in the reference; counted, and zero
for all sets here). Split-precision path: hit ids must agree on every ray that is neither a near-tie nor
order-dependent (two cylinders accepted), and agreeing hits must have identical t/u/v (|dt|/t <= 1e-12 bar).
"""
import json
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401  (raises if libblazert_amd.so is missing)
    return torch.device("cuda:0")


def gpu_intersect(scene, o, d, dev, tmin=None, tmax=None, precision=0, any_hit=False, normals=True, **launch):
    O = torch.from_numpy(np.ascontiguousarray(o)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    kw = {}
    if tmin is not None:
        kw["tmin"] = torch.from_numpy(np.ascontiguousarray(tmin, dtype=np.float64)).to(dev)
    if tmax is not None:
        kw["tmax"] = torch.from_numpy(np.ascontiguousarray(tmax, dtype=np.float64)).to(dev)
    h = scene.intersect(O, D, precision=precision, any_hit=any_hit, normals=normals, **kw, **launch)
    torch.cuda.synchronize()
    out = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
    out["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
    out["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
    if normals:
        out["normal"] = h["normal"].cpu().numpy().T
    return out


def make_pair(oracle, fill, **opts):
    from blazert_amd import Scene
    p = Scene(**opts)
    fill(p)
    p.commit()
    o = oracle.Scene(*(opts.get(k, dflt) for k, dflt in (("min_leaf_primitives", 4), ("max_tree_depth", 28),
                                                           ("bin_size", 64))))
    fill(o)
    o.commit()
    return p, o


def assert_bit_exact(g, r, mask=None, normals=True):
    m = np.ones(len(g["t"]), bool) if mask is None else mask
    fields = ["t", "u", "v", "prim_id", "geom_id"]
    for k in fields:
        a, b = np.asarray(g[k])[m], np.asarray(r[k])[m]
        bad = np.nonzero(~_same(a, b))[0]
        assert len(bad) == 0, f"{k}: {len(bad)} mismatches, first rays {np.nonzero(m)[0][bad[:5]]}"
    if normals:
        same = _same(np.asarray(g["normal"])[m], np.asarray(r["normal"])[m])
        assert same.all(), f"normal: {int((~same).any(axis=1).sum())} mismatches"


def _same(a, b):
    """Bitwise equality; two NaNs count as equal (their payload bits differ between CPU and GPU)."""
    if a.dtype != np.float64:
        return a == b
    return (a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))


# ------------------------------------------------------------------------------------------------ known answers
DOC = json.load(open(os.path.join(GOLDEN, "known_answers.json")))


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


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("case", [pytest.param(c, id=c["name"]) for c in DOC["cases"]
                                  if c["kind"] in ("primitive", "scene", "traverse", "traverse_distance")])
def test_known_answers_on_gpu(dev, oracle_mod, case, precision):
    """Both precisions give the oracle's bits: these rays have no ties, and a split-precision cylinder hit here has
    no second cylinder to disagree with."""
    p, o = make_pair(oracle_mod, _fill_case(case))
    org = np.array([case["org"]], float)
    dirs = np.array([case["dir"]], float)
    g = gpu_intersect(p, org, dirs, dev, precision=precision)
    r = o.intersect_batch(org, dirs)
    assert_bit_exact(g, r)
    assert (g["geom_id"][0] != INVALID) == case["hit"]
    if case["hit"]:
        eps = DOC["approx_epsilon"]
        assert abs(g["t"][0] - case["t"]) < eps * (1 + abs(case["t"]))


def test_scene_primitives_golden_image(dev, oracle_mod):
    """examples/baseline/scene_primitives.png reproduced pixel for pixel (cylinder quirks included)."""
    sc = W.scene_primitives()
    p, o = make_pair(oracle_mod, lambda s: (s.add_cylinders(*sc["cylinders"]), s.add_spheres(*sc["spheres"])))
    org, dirs = W.scene_primitives_rays()
    g = gpu_intersect(p, org, dirs, dev)
    labels = np.load(os.path.join(GOLDEN, "scene_primitives_labels.npz"))["labels"]
    mine = np.where(g["geom_id"] == INVALID, 0, np.where(g["prim_id"] == 0, 1, 2))
    assert int((mine != labels).sum()) == 0
    assert_bit_exact(g, o.intersect_batch(org, dirs))


def test_sphere_field_full(dev, oracle_mod):
    """1,024 spheres, 512x512 pinhole rays, tmax = +inf."""
    c, r = W.sphere_field()
    p, o = make_pair(oracle_mod, lambda s: s.add_spheres(c, r))
    org, dirs = W.sphere_field_rays()
    tmax = np.full(len(org), np.inf)
    g = gpu_intersect(p, org, dirs, dev, tmax=tmax)
    ref = o.intersect_batch(org, dirs, tmax=tmax, trace=True)
    assert ref["trace"]["overflow"].sum() == 0
    assert_bit_exact(g, ref)
    assert (g["geom_id"] != INVALID).mean() > 0.05


@pytest.fixture(scope="module")
def ico8(oracle_mod):
    v, f = W.icosphere(8)
    return make_pair(oracle_mod, lambda s: s.add_mesh(v, f))


def test_icosphere8_incoherent_subset(dev, ico8):
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays(1 << 18)
    g = gpu_intersect(p, org, dirs, dev)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    assert ok.mean() > 0.999
    assert_bit_exact(g, ref, ok)


@pytest.mark.parametrize("sort", [1, 2])
@pytest.mark.parametrize("precision", [0, 1])
def test_icosphere8_sorted_launch(dev, ico8, sort, precision):
    """Sorted launches trace rays in another order but write each hit to its own ray: same bits as the oracle."""
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays(1 << 18)
    g = gpu_intersect(p, org, dirs, dev, precision=precision, sort=sort)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = (ref["trace"]["overflow"] == 0) & ((ref["trace"]["tie"] == 0) | (precision == 0))
    assert_bit_exact(g, ref, ok)


@pytest.mark.parametrize("schedule,refill,block,ring", [(1, 0, 64, 0), (0, 1, 64, 4), (0, 64, 128, 16),
                                                        (0, 16, 192, 8), (1, 0, 256, 4)])
def test_launch_configurations_identical(dev, ico8, schedule, refill, block, ring):
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays(1 << 16, seed=21)
    g = gpu_intersect(p, org, dirs, dev, schedule=schedule, refill=refill, block_size=block, stack_ring=ring)
    assert_bit_exact(g, o.intersect_batch(org, dirs))
    g32 = gpu_intersect(p, org, dirs, dev, precision=1, schedule=schedule, refill=refill, block_size=block,
                        stack_ring=ring)
    base = gpu_intersect(p, org, dirs, dev, precision=1)
    assert_bit_exact(g32, base)  # the split path is deterministic per ray, whatever the launch shape


@pytest.mark.parametrize("pool", [0, 32, 128])
def test_ragged_ray_counts_claim_every_ray_once(dev, ico8, pool):
    """Ray counts around the claim boundaries: fewer rays than one claim per counter (so some of the eight claim
    counters never hand out a chunk), one past and one short of chunk multiples, and a last chunk of one ray.
    Every ray must be traced exactly once: same bits as the oracle on every ray."""
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays(8 * 128 * 3 + 1, seed=23)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    for n in (1, 31, 33, 65, 257, 511, 513, 8 * 32 + 1, 8 * 64 - 1, 8 * 128 + 1, 8 * 128 * 3 + 1):
        g = gpu_intersect(p, org[:n], dirs[:n], dev, pool_size=pool, normals=False)
        assert_bit_exact(g, {k: ref[k][:n] for k in ("t", "u", "v", "prim_id", "geom_id")}, ok[:n], normals=False)
        g = gpu_intersect(p, org[:n], dirs[:n], dev, precision=1, pool_size=pool, normals=False)
        h = gpu_intersect(p, org[:n], dirs[:n], dev, precision=1, schedule=1, normals=False)
        assert_bit_exact(g, h, normals=False)
    from blazert_amd.capi import BlazertError
    with pytest.raises(BlazertError, match="pool_size"):
        gpu_intersect(p, org[:64], dirs[:64], dev, pool_size=96)


def test_icosphere8_coherent_subset(dev, ico8):
    p, o = ico8
    org, dirs = W.icosphere_coherent_rays(640, 360)
    g = gpu_intersect(p, org, dirs, dev)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    assert_bit_exact(g, ref, ok)


def test_icosphere8_full_persistent_equals_per_thread(dev, ico8):
    """All 4,194,304 incoherent rays, every field: the persistent schedule (wave-level ray pools claimed from eight
    counters, lane refill) against one ray per thread, at both precisions. A ray skipped or traced twice by the pool
    claims would show here."""
    p, _ = ico8
    org, dirs = W.icosphere_incoherent_rays()
    for precision in (0, 1):
        a = gpu_intersect(p, org, dirs, dev, precision=precision)
        b = gpu_intersect(p, org, dirs, dev, precision=precision, schedule=1)
        assert_bit_exact(a, b)
        # 32- and 128-index claims (the pool_size launch option) hand out the same rays
        for pool in (32, 128):
            c = gpu_intersect(p, org, dirs, dev, precision=precision, pool_size=pool)
            assert_bit_exact(a, c)


def test_icosphere8_full_incoherent_properties(dev, ico8):
    """All 4,194,304 incoherent rays: size-independent properties (the subset above is compared bit for bit)."""
    p, _ = ico8
    org, dirs = W.icosphere_incoherent_rays()
    g = gpu_intersect(p, org, dirs, dev, normals=False)
    hit = g["geom_id"] != INVALID
    assert 0.05 < hit.mean() < 0.9
    assert (g["t"][hit] > 0).all() and (g["t"][~hit] == np.finfo(np.float64).max).all()
    assert (g["prim_id"][hit] < 1310720).all()
    u, v = g["u"][hit], g["v"][hit]
    tol = 1e-9
    assert (u >= -tol).all() and (v >= -tol).all() and (u + v <= 1 + tol).all()
    # tracing a prefix of the set as its own batch gives the same bits (batch-size independence)
    g2 = gpu_intersect(p, org[:1 << 18], dirs[:1 << 18], dev, normals=False)
    assert np.array_equal(g2["t"].view(np.uint64), g["t"][:1 << 18].view(np.uint64))


@pytest.fixture(scope="module")
def mixed(oracle_mod):
    m = W.mixed_scene(n_tri_target=40000, n_sph=10000, n_cyl=2000, n_pla=100, seed=17)

    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    return make_pair(oracle_mod, fill)


def test_mixed_fp64_exact(dev, mixed):
    p, o = mixed
    org, dirs = W.mixed_rays(1 << 18, seed=5)
    g = gpu_intersect(p, org, dirs, dev)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    assert_bit_exact(g, ref, ok)
    assert len(np.unique(g["geom_id"])) == 5  # four geometry types and misses all present


def test_mixed_split_precision_agreement(dev, mixed, oracle_mod):
    """Every disagreement of the split path with the FP64 reference is a tie or a cylinder case (the classes the split
    path's error bound allows); nothing else."""
    p, o = mixed
    org, dirs = W.mixed_rays(1 << 18, seed=5)
    g32 = gpu_intersect(p, org, dirs, dev, precision=1)
    ref = o.intersect_batch(org, dirs)
    same = (g32["geom_id"] == ref["geom_id"]) & (g32["prim_id"] == ref["prim_id"])
    idx = np.nonzero(~same)[0]
    cls = oracle_mod.classify_split_disagreements(o, org, dirs, idx, g32["t"][idx], g32["geom_id"][idx], cyl_geom=2)
    assert cls["unexplained"] == 0, (cls, cls["unexplained_idx"][:10])
    assert_bit_exact(g32, ref, same)  # every agreeing ray: the same t, u, v and normal bits


def test_icosphere_split_precision_bit_identical(dev, ico8):
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays(1 << 18)
    g32 = gpu_intersect(p, org, dirs, dev, precision=1)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = (ref["trace"]["overflow"] == 0) & (ref["trace"]["tie"] == 0)
    assert_bit_exact(g32, ref, ok)


def test_any_hit_matches_oracle(dev, mixed):
    p, o = mixed
    org, dirs = W.mixed_rays(1 << 16, seed=9)
    g = gpu_intersect(p, org, dirs, dev, any_hit=True)
    ref = o.intersect_batch(org, dirs, trace=True, any_hit=True)
    assert_bit_exact(g, ref, ref["trace"]["overflow"] == 0)


def test_inverted_boxes_take_the_exact_general_path(dev, oracle_mod):
    """Spheres with a negative radius have inverted boxes (center - r > center + r). The slab test's min / max form
    is exact only for boxes with min <= max, so the host flags such trees and the kernel takes the sign-selected
    general form: hits must still equal the oracle bit for bit, at both precisions."""
    c, r = W.sphere_field(4096, seed=41)
    r = np.where(W.Rng(42).uniform(len(r)) < 0.3, -r, r)
    p, o = make_pair(oracle_mod, lambda s: s.add_spheres(c, r))
    org, dirs = W.incoherent_rays(1 << 15, 43, 12.0)
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    g = gpu_intersect(p, org, dirs, dev)
    assert (ref["prim_id"] != INVALID).sum() > 1000
    assert_bit_exact(g, ref, ok)
    # the split path's FP32 boxes bound what the tree's boxes bound; an inverted box bounds nothing, so such a tree
    # is traced by the FP64 path whatever the requested precision
    g32 = gpu_intersect(p, org, dirs, dev, precision=1)
    assert_bit_exact(g32, g)
    # the same scene built by the device builder (finite boxes: it builds them; the host flags the order)
    from blazert_amd import Scene
    d = Scene()
    d.add_spheres(c, r)
    d.commit(device=True)
    assert d.build_stats(1)["built_on_device"] == 1
    assert_bit_exact(gpu_intersect(d, org, dirs, dev), ref, ok)
    # a root that is a leaf: one or two spheres with negative radii, so the root box itself is inverted
    for cs, rs in (([[0.0, 0.0, 0.0]], [-1.0]), ([[0.0, 0.0, 0.0], [0.5, 0.0, 0.0]], [-1.0, -0.75])):
        p1, o1 = make_pair(oracle_mod, lambda s: s.add_spheres(np.array(cs), np.array(rs)))
        org1, dirs1 = W.incoherent_rays(4096, 45, 3.0)
        ref1 = o1.intersect_batch(org1, dirs1)
        assert_bit_exact(gpu_intersect(p1, org1, dirs1, dev), ref1)


@pytest.mark.parametrize("precision", [0, 1])
def test_degenerate_and_axis_parallel_rays(dev, oracle_mod, precision):
    """Zero direction components (1/d = inf), origins on box planes (0*inf = NaN slab terms), -0.0, tmin/tmax.
    The split path runs here where its error bound is tightest: |1/d| = inf axes (never culled), boxes touching the
    origin, and tmin windows; any disagreement with the FP64 reference must be a tie or a cylinder case."""
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1], [-1, 0, 0], [0, -1, 0]], float)
    f = np.array([[0, 1, 2], [0, 2, 3], [0, 3, 1], [1, 4, 2], [0, 5, 6], [4, 3, 1], [2, 4, 3]], np.uint32)
    cen = np.array([[0.5, 0.5, 0.5], [2, 2, 2], [0, 0, 0]])
    rad = np.array([0.25, 1.0, 0.5])

    def fill(s):
        s.add_mesh(v, f)
        s.add_spheres(cen, rad)
        s.add_planes(np.array([[0, 0, 0]], float), [2.0], [2.0], np.eye(3).reshape(1, 9))
        s.add_cylinders(np.array([[0, 0, 1.0]]), [0.5], [0.5], [1.0], np.eye(3).reshape(1, 9))

    p, o = make_pair(oracle_mod, fill)
    vals = np.array([-1.0, -0.5, -0.0, 0.0, 0.25, 0.5, 1.0, 2.0])
    r = W.Rng(3)
    n = 4096
    org = vals[(r.uniform(3 * n) * len(vals)).astype(int)].reshape(n, 3)
    dirs = np.array([-1.0, -0.0, 0.0, 1.0, 0.5])[(r.uniform(3 * n) * 5).astype(int)].reshape(n, 3)
    tmin = np.where(r.uniform(n) < 0.2, 0.3, 0.0)
    tmax = np.where(r.uniform(n) < 0.2, 1.5, np.finfo(np.float64).max)
    g = gpu_intersect(p, org, dirs, dev, tmin=tmin, tmax=tmax, precision=precision)
    ref = o.intersect_batch(org, dirs, tmin=tmin, tmax=tmax)
    # rays with an all-zero direction normalise to 0 and have 1/d = inf on every axis; they are compared too
    if precision == 0:
        assert_bit_exact(g, ref)
        return
    same = (g["geom_id"] == ref["geom_id"]) & (g["prim_id"] == ref["prim_id"])
    idx = np.nonzero(~same)[0]
    cls = oracle_mod.classify_split_disagreements(o, org, dirs, idx, g["t"][idx], g["geom_id"][idx], cyl_geom=3,
                                                  tmin=tmin, tmax=tmax)
    assert cls["unexplained"] == 0, (cls, cls["unexplained_idx"][:10])
    assert_bit_exact(g, ref, same)
    assert same.mean() > 0.99
    # rays with a zero direction component are deferred to the FP64 path: bit-exact to the reference
    zero_dir = (dirs == 0.0).any(axis=1)
    assert zero_dir.sum() > 1000
    assert_bit_exact(g, ref, zero_dir)
    # the same rays as a one-item split batch, captured and replayed (the deferred FP64 launch is part of the graph)
    from blazert_amd import Batch
    col = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    b = Batch([(p, col(org), col(dirs), col(tmin), col(tmax))], precision=1)
    s = torch.cuda.Stream()
    b.capture(s)
    for _ in range(2):
        b.hits[0]["t"].fill_(-1.0)
        b.hits[0]["prim_id"].fill_(-7)
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        h = b.hits[0]
        assert np.array_equal(h["t"].cpu().numpy().view(np.uint64), g["t"].view(np.uint64))
        assert np.array_equal(h["prim_id"].cpu().numpy().view(np.uint32), g["prim_id"])
        assert np.array_equal(h["geom_id"].cpu().numpy().view(np.uint32), g["geom_id"])


def test_hits_inout_semantics(dev, oracle_mod):
    """RayHit in/out: a prior hit distance is the merge threshold for non-triangle types; misses keep it."""
    c, r = W.sphere_field(64, seed=4)
    p, _ = make_pair(oracle_mod, lambda s: s.add_spheres(c, r))
    org, dirs = W.camera_rays(64, 64, eye=(0, 0, 30))
    O = torch.from_numpy(org).to(dev)
    D = torch.from_numpy(dirs).to(dev)
    base = p.intersect(O, D)
    torch.cuda.synchronize()
    prior = {k: v.clone() for k, v in base.items()}
    prior["t"].fill_(5.0)
    prior["prim_id"].fill_(123)
    prior["geom_id"].fill_(7)
    out = p.intersect(O, D, hits=prior)
    torch.cuda.synchronize()
    t0 = base["t"].cpu().numpy()
    closer = t0 < 5.0
    assert np.array_equal(out["t"].cpu().numpy()[closer], t0[closer])
    assert (out["t"].cpu().numpy()[~closer] == 5.0).all()
    assert (out["prim_id"].cpu().numpy()[~closer] == 123).all()


def test_intersect1_dropin(dev, oracle_mod):
    from blazert_amd import Ray, RayHit, intersect1
    sc = W.scene_primitives()
    p, o = make_pair(oracle_mod, lambda s: (s.add_cylinders(*sc["cylinders"]), s.add_spheres(*sc["spheres"])))
    for org, d in [((0.0, 5.0, 20.0), (0.0, -0.2, -1.0)), ((2.0, 1.0, 0.0), (-1.0, 0.0, 0.0)),
                   ((0.0, 5.0, 20.0), (0.3, 0.3, -1.0))]:
        rh = RayHit()
        hit = intersect1(p, Ray(org, d), rh)
        ok, ref, _ = o.intersect1(oracle_mod.make_ray(org, d))
        assert hit == ok
        assert rh.prim_id == ref.prim_id and rh.geom_id == ref.geom_id
        assert np.float64(rh.hit_distance).view(np.uint64) == np.float64(ref.t).view(np.uint64)


def test_empty_and_single_prim_scenes(dev, oracle_mod):
    from blazert_amd import Scene
    s = Scene()
    s.commit()
    org, dirs = W.camera_rays(8, 8, eye=(0, 0, 3))
    g = gpu_intersect(s, org, dirs, dev)
    assert (g["geom_id"] == INVALID).all() and (g["t"] == np.finfo(np.float64).max).all()
    p, o = make_pair(oracle_mod, lambda x: x.add_spheres(np.zeros((1, 3)), [1.0]))
    assert_bit_exact(gpu_intersect(p, org, dirs, dev), o.intersect_batch(org, dirs))


# ------------------------------------------------------------------------------------------------ batches
@pytest.fixture(scope="module")
def batch_set(oracle_mod):
    """Eight uneven scenes (icospheres, sphere fields and a mixed scene) with uneven ray batches."""
    from blazert_amd import Scene
    rng = W.Rng(77)
    items = []
    for i in range(8):
        kind = i % 3
        if kind == 0:
            v, f = W.icosphere(3 + i % 4, seed=i)
            fill = (lambda v, f: lambda s: s.add_mesh(v * 2.0, f))(v, f)
        elif kind == 1:
            c, r = W.sphere_field(200 + 300 * i, seed=i)
            fill = (lambda c, r: lambda s: s.add_spheres(c * 0.2, r * 0.2))(c, r)
        else:
            m = W.mixed_scene(n_tri_target=3000, n_sph=500, n_cyl=100, n_pla=10, seed=i)
            fill = (lambda m: lambda s: (s.add_mesh(*m["mesh"]), s.add_spheres(*m["spheres"]),
                                         s.add_cylinders(*m["cylinders"]), s.add_planes(*m["planes"])))(m)
        p = Scene()
        fill(p)
        p.commit()
        o = oracle_mod.Scene()
        fill(o)
        o.commit()
        n = int(2 ** (10 + 6 * rng.uniform(1)[0]))
        org, dirs = W.incoherent_rays(n, seed=100 + i, box=4.0 if kind != 2 else 22.0)
        items.append((p, o, org, dirs))
    empty = Scene()
    empty.commit()
    org, dirs = W.incoherent_rays(777, seed=5, box=2.0)
    items.append((empty, None, org, dirs))
    return items


def _batch_check(batch, items, precision):
    for (p, o, org, dirs), h in zip(items, batch.hits):
        g = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
        g["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
        g["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
        if o is None:
            assert (g["geom_id"] == INVALID).all()
            continue
        ref = o.intersect_batch(org, dirs, trace=True)
        tr = ref["trace"]
        ok = tr["overflow"] == 0
        if precision:
            ok &= (tr["tie"] == 0) & (tr["multi_cyl"] == 0)
        assert_bit_exact(g, ref, ok, normals=False)


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_launch_matches_oracle(dev, batch_set, precision):
    from blazert_amd import Batch
    items = [(p, torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)) for p, o, org, dirs in batch_set]
    b = Batch(items, precision=precision)
    b.launch()
    torch.cuda.synchronize()
    _batch_check(b, batch_set, precision)


def test_batch_graph_replays(dev, batch_set):
    """A captured hipGraph replayed several times, with eager work in between, gives the eager result each time."""
    from blazert_amd import Batch
    items = [(p, torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)) for p, o, org, dirs in batch_set]
    b = Batch(items, precision=1)
    s = torch.cuda.Stream()
    b.capture(s)
    for rep in range(3):
        for h in b.hits:
            h["t"].fill_(-1.0)
            h["prim_id"].fill_(-7)
        torch.cuda.synchronize()
        b.replay(s)
        s.synchronize()
        _batch_check(b, batch_set, 1)


@pytest.mark.parametrize("njobs", [32, 33])
def test_batch_many_jobs_equals_single_launches(dev, njobs):
    """Batches of 32 jobs (the kernel keeps each job's hit columns in LDS) and 33 (it reads them from the job
    table): every field of every ray equals the scene's own intersect, at both precisions, and rays of the
    spheres that follow the triangles merge into the triangle hits."""
    from blazert_amd import Batch, Scene
    items = []
    for i in range(njobs):
        v, f = W.icosphere(2, seed=i)
        s = Scene()
        s.add_mesh(v * (1.0 + 0.1 * (i % 5)), f)
        if i % 3 == 0:
            c, r = W.sphere_field(50, seed=i)
            s.add_spheres(c * 0.1, r * 0.1)
        s.commit()
        org, dirs = W.incoherent_rays(300 + 97 * i, seed=200 + i, box=3.0)
        items.append((s, torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)))
    for precision in (0, 1):
        b = Batch(items, precision=precision)
        b.launch()
        torch.cuda.synchronize()
        for (s, O, D), h in zip(items, b.hits):
            ref = s.intersect(O, D, precision=precision)
            torch.cuda.synchronize()
            for k in ("t", "u", "v", "prim_id", "geom_id"):
                assert torch.equal(h[k], ref[k]), (precision, k)


def test_low_level_traverse_per_type(dev, mixed, oracle_mod):
    """brt_traverse (blazert/bvh/accel.h:60-120) on each geometry type: hit rays get the type's post_traversal
    record bit for bit, missed rays keep the record passed in."""
    p, o = mixed
    org, dirs = W.mixed_rays(4096, seed=9)
    O, D = torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)
    geom_of_type = {0: 0, 1: 1, 3: 2, 2: 3}  # add order of the fixture: mesh, spheres, cylinders, planes
    for typ in range(4):
        n = len(org)
        hits = {"t": torch.full((n,), -1.0, dtype=torch.float64, device=dev),
                "u": torch.full((n,), -2.0, dtype=torch.float64, device=dev),
                "v": torch.full((n,), -2.0, dtype=torch.float64, device=dev),
                "prim_id": torch.full((n,), 123456, dtype=torch.int32, device=dev),
                "geom_id": torch.full((n,), 7, dtype=torch.int32, device=dev),
                "normal": torch.full((3, n), -3.0, dtype=torch.float64, device=dev)}
        p.traverse(typ, O, D, hits)
        torch.cuda.synchronize()
        g = {k: hits[k].cpu().numpy() for k in ("t", "u", "v")}
        gp = hits["prim_id"].cpu().numpy().view(np.uint32)
        gg = hits["geom_id"].cpu().numpy().view(np.uint32)
        gn = hits["normal"].cpu().numpy().T
        nhit = 0
        for i in range(n):
            ok, h, tr = o.traverse(typ, oracle_mod.make_ray(org[i], dirs[i]))
            if tr.overflow:
                continue
            if ok:
                nhit += 1
                assert np.float64(g["t"][i]).view(np.uint64) == np.float64(h.t).view(np.uint64), (typ, i)
                assert np.float64(g["u"][i]).view(np.uint64) == np.float64(h.u).view(np.uint64), (typ, i)
                assert np.float64(g["v"][i]).view(np.uint64) == np.float64(h.v).view(np.uint64), (typ, i)
                assert gp[i] == h.prim_id and gg[i] == geom_of_type[typ], (typ, i)
                assert np.array_equal(gn[i].view(np.uint64), np.array(h.normal[:]).view(np.uint64)), (typ, i)
            else:
                assert g["t"][i] == -1.0 and gp[i] == 123456 and gg[i] == 7 and (gn[i] == -3.0).all(), (typ, i)
        assert nhit > 0, typ


# ------------------------------------------------------------------------------------------------ brute force
def test_fp64_equals_bruteforce_icosphere8(dev, ico8, oracle_mod):
    """The GPU's FP64 hits against a brute-force search (every triangle tested, no tree) on a sample of the
    1,310,720-triangle incoherent set: the tree may only reorder ties. The reference's traversal stack stays within
    its 28 slots on every ray of this set (checked), so no ray is excluded for overflow."""
    p, o = ico8
    org, dirs = W.icosphere_incoherent_rays()
    sample = np.arange(0, len(org), len(org) // 384)
    g = gpu_intersect(p, org[sample], dirs[sample], dev, normals=False)
    ref = o.intersect_batch(org[sample], dirs[sample], normals=False, trace=True)
    assert (ref["trace"]["overflow"] == 0).all()
    bt, bp, bg = o.bruteforce_merged(org[sample], dirs[sample])
    ok = ref["trace"]["tie"] == 0
    assert ok.mean() > 0.99 and (bp != INVALID).sum() > 50
    assert np.array_equal(g["t"][ok].view(np.uint64), bt[ok].view(np.uint64))
    assert np.array_equal(g["prim_id"][ok], bp[ok]) and np.array_equal(g["geom_id"][ok], bg[ok])


def test_fp64_equals_bruteforce_mixed_scene(dev, oracle_mod):
    """Mixed scene (200k-triangle target, 50k spheres, 10k cylinders, 100 planes), incoherent rays: the GPU's FP64
    hits against the brute-force closest triangle / sphere / plane merged in the reference's type order, on every
    sampled ray that no cylinder accepts (a cylinder is accepted without a distance check, so the last visited wins
    in the reference and brute force has no order to follow) and that is not a near-tie."""
    m = W.mixed_scene(seed=11)

    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    p, o = make_pair(oracle_mod, fill)
    org, dirs = W.mixed_rays(3072, seed=23)
    g = gpu_intersect(p, org, dirs, dev, normals=False)
    ref = o.intersect_batch(org, dirs, normals=False, trace=True)
    bt, bp, bg = o.bruteforce_merged(org, dirs)
    _, cyl = o.bruteforce_batch(3, org, dirs)
    ok = (ref["trace"]["tie"] == 0) & (ref["trace"]["overflow"] == 0) & (cyl == INVALID)
    assert ok.sum() > 1000 and (bp[ok] != INVALID).sum() > 300
    assert np.array_equal(g["t"][ok].view(np.uint64), bt[ok].view(np.uint64))
    assert np.array_equal(g["prim_id"][ok], bp[ok]) and np.array_equal(g["geom_id"][ok], bg[ok])


def test_reference_stack_overflow_rays_equal_bruteforce(dev, oracle_mod):
    """Rays on which the reference's 28-slot traversal stack overflows (undefined behaviour there). A deep tree:
    120 unit spheres at x = 100 * 1.3^-k plus one large sphere beyond them, built with leaf size 1, depth limit 60
    and 2 bins, so each split peels off the outermost sphere. Rays along +x pass through every box corner and miss
    the unit spheres, so the traversal descends the whole chain pushing each far child. The GPU's deeper stack must
    give the brute-force closest hit on every such ray."""
    k = np.arange(120, dtype=float)
    x = 100.0 * 1.3 ** (-k)
    c = np.concatenate([np.stack([x, 0 * x, 0 * x], 1), [[150.0, 0.0, 0.0]]])
    r = np.concatenate([np.full(120, 0.5), [5.0]])
    p, o = make_pair(oracle_mod, lambda s: s.add_spheres(c, r), min_leaf_primitives=1, max_tree_depth=60, bin_size=2)
    rng = W.Rng(5)
    n = 4096
    org = np.stack([np.full(n, -10.0), rng.uniform(n, 0.3, 0.49), rng.uniform(n, 0.3, 0.49)], 1)
    dirs = np.stack([np.full(n, 1.0), rng.uniform(n, -2e-3, 2e-3), rng.uniform(n, -2e-3, 2e-3)], 1)
    ref = o.intersect_batch(org, dirs, trace=True)
    ovf = ref["trace"]["overflow"] != 0
    assert ovf.sum() > 3000
    bt, bp, bg = o.bruteforce_merged(org, dirs, types=(1,))
    assert (bp[ovf] == 120).sum() > 1000 and (bp[ovf] != 120).sum() > 100
    for precision in (0, 1):
        g = gpu_intersect(p, org, dirs, dev, normals=False, precision=precision)
        assert np.array_equal(g["t"].view(np.uint64), bt.view(np.uint64)), precision
        assert np.array_equal(g["prim_id"], bp) and np.array_equal(g["geom_id"], bg), precision


# ------------------------------------------------------------------------------------------------ uneven batch
def test_uneven_batch_32_scenes(dev, oracle_mod):
    """The batched workload at its full shape: 32 triangle scenes (displaced icospheres and tori, 1k-500k triangles,
    log-uniform) with 16k-2M incoherent rays each (15.6M rays in all), split precision. Leaf size is tuned per
    scene, block size / refill / LDS stack depth for the batch; the batch is captured in a hipGraph and replayed
    twice. A sample of 2,048 rays per scene is compared with the FP64 oracle built with the tuned leaf size: every
    disagreement must be a tie (triangles only: no cylinder class), every agreeing ray bit-identical in t, u, v."""
    from blazert_amd import PRECISION_FP32_TRAV_FP64_HIT, Batch, Scene
    from blazert_amd.autotune import tune_batch, tune_leaf

    specs = W.uneven_batch()
    assert len(specs) == 32
    items, oracles = [], []
    for sp in specs:
        s = Scene()
        s.add_mesh(sp["vertices"], sp["faces"])
        O = torch.from_numpy(np.ascontiguousarray(sp["origins"])).to(dev)
        D = torch.from_numpy(np.ascontiguousarray(sp["directions"])).to(dev)
        m = min(len(sp["origins"]), 1 << 14)
        leaf = tune_leaf(s, O[:m], D[:m], precision=PRECISION_FP32_TRAV_FP64_HIT, leaf_sizes=(2, 4, 8),
                         repeats=1)["min_leaf_primitives"]
        o = oracle_mod.Scene(leaf, 28, 64)
        o.add_mesh(sp["vertices"], sp["faces"])
        o.commit()
        items.append((s, O, D))
        oracles.append(o)
    b = Batch(items, precision=PRECISION_FP32_TRAV_FP64_HIT)
    tune_batch(b, repeats=1)
    stream = torch.cuda.Stream()
    b.capture(stream)
    for rep in range(2):
        for h in b.hits:
            h["t"].fill_(-1.0)
            h["prim_id"].fill_(-7)
        torch.cuda.synchronize()
        b.replay(stream)
        stream.synchronize()
        for k, (sp, o, h) in enumerate(zip(specs, oracles, b.hits)):
            n = len(sp["origins"])
            idx = np.linspace(0, n - 1, 2048).astype(np.int64)
            org, dirs = sp["origins"][idx], sp["directions"][idx]
            g = {key: h[key].cpu().numpy()[idx] for key in ("t", "u", "v")}
            g["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)[idx]
            g["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)[idx]
            ref = o.intersect_batch(org, dirs)
            same = (g["prim_id"] == ref["prim_id"]) & (g["geom_id"] == ref["geom_id"])
            bad = np.nonzero(~same)[0]
            cls = oracle_mod.classify_split_disagreements(o, org, dirs, bad, g["t"][bad], g["geom_id"][bad])
            assert cls["unexplained"] == 0, (k, cls)
            assert_bit_exact(g, ref, same, normals=False)
            assert (g["geom_id"] != INVALID).sum() > 100, k


def test_side_stream_equals_default_stream(dev, ico8):
    """Launches on a side stream: the library waits for the current stream's temporaries (the columns copied out of
    the (n,3) tensors, a broadcast tmax) and keeps them and the new hit tensors alive for the side stream's kernels.
    The result must be bit-identical to the default stream's while the caching allocator recycles memory between
    launches."""
    p, _ = ico8
    org, dirs = W.icosphere_incoherent_rays(1 << 18, seed=31)
    ref = gpu_intersect(p, org, dirs, dev, tmax=np.full(len(org), 3.0))
    side = torch.cuda.Stream(device=dev)
    O = torch.from_numpy(org).to(dev)
    D = torch.from_numpy(dirs).to(dev)
    for rep in range(3):
        h = p.intersect(O, D, tmax=3.0, normals=True, stream=side)
        junk = [torch.empty(1 << 20, dtype=torch.float64, device=dev).fill_(rep) for _ in range(8)]  # noqa: F841
        side.synchronize()
        g = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
        g["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
        g["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
        g["normal"] = h["normal"].cpu().numpy().T
        assert_bit_exact(g, ref)
