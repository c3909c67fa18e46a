"""Device build (brt_scene_commit_device) against the host builder, through the C ABI.

The host builder is pinned to the reference's tree by the oracle tests; the device build must reproduce it byte for
byte: the reference-form tree (every node's box, axis, children and leaf range), the leaf order, the build
statistics and all four device arrays (FP64 nodes, FP32 4-wide nodes, primitive records, primitive ids). Cases
cover both builder regimes (level-synchronous large nodes, one-wave small subtrees), every geometry type, every
build option, degenerate splits, depth-limited leaves, tiny scenes, the fallback to the host builder, rebuilds by
the leaf-size tuner, and the 1,310,720-triangle icosphere.
"""
import time

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import capi  # noqa: E402
from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

ARRAYS = (capi.ARRAY_NODES64, capi.ARRAY_NODES4, capi.ARRAY_PRIMITIVES, capi.ARRAY_PRIM_IDS)
STATS = ("max_tree_depth", "leaf_nodes", "max_num_prims_per_leaf", "min_num_prims_per_leaf", "branch_nodes",
         "bad_splits", "num_primitives", "geom_id")


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401
    return torch.device("cuda:0")


def _pair(fill, **opts):
    from blazert_amd import Scene
    h = Scene(**opts)
    fill(h)
    h.commit()
    d = Scene(**opts)
    fill(d)
    d.commit(device=True)
    torch.cuda.synchronize()
    return h, d


def _assert_same_build(h, d, types, on_device=True):
    for t in types:
        sh, sd = h.build_stats(t), d.build_stats(t)
        for k in STATS:
            assert sh[k] == sd[k], (t, k, sh[k], sd[k])
        assert sh["built_on_device"] == 0
        if on_device is not None:
            assert sd["built_on_device"] == int(on_device), t
        nh, oh = h.export_tree(t)
        nd, od = d.export_tree(t)
        assert len(nh) == len(nd), t
        assert nh.tobytes() == nd.tobytes(), (t, "tree")
        assert np.array_equal(oh, od), (t, "order")
        for a in ARRAYS:
            x, y = h.read_device_array(t, a), d.read_device_array(t, a)
            assert x.size == y.size and np.array_equal(x, y), (t, a, x.size, y.size)


def _mesh(V, F):
    return lambda s: s.add_mesh(V, F)


@pytest.mark.parametrize("level", [0, 1, 3, 5, 7])
def test_icosphere_levels(dev, level):
    V, F = W.icosphere(level)
    print(f"device build: icosphere level {level}, {len(F)} triangles", flush=True)
    _assert_same_build(*_pair(_mesh(V, F)), [0])


@pytest.mark.parametrize("opts", [dict(min_leaf_primitives=1), dict(min_leaf_primitives=2),
                                  dict(min_leaf_primitives=8), dict(bin_size=2), dict(bin_size=16),
                                  dict(bin_size=33), dict(max_tree_depth=3), dict(max_tree_depth=10),
                                  dict(min_leaf_primitives=300, bin_size=7)])
def test_build_options(dev, opts):
    V, F = W.icosphere(5)
    print(f"device build: icosphere level 5, options {opts}", flush=True)
    _assert_same_build(*_pair(_mesh(V, F), **opts), [0])


def _fill_mixed(m):
    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    return fill


def test_mixed_scene_all_types(dev):
    m = W.mixed_scene(20_000, 5_000, 1_000, 100, seed=3)
    print("device build: mixed scene", flush=True)
    _assert_same_build(*_pair(_fill_mixed(m)), [0, 1, 2, 3])


def test_degenerate_splits(dev):
    """Identical centers force every partition attempt to fail (bad splits, middle fallback, depth limit)."""
    r = W.Rng(5)
    n = 3000
    c = np.zeros((n, 3)) + np.array([1.5, -2.0, 0.25])
    rad = r.uniform(n, 0.1, 1.0)
    # a mesh of many copies of two triangles plus a spread-out part
    base = np.array([[0.0, 0.0, 1.0], [1.0, 0.0, 1.0], [0.0, 1.0, 1.0], [2.0, 2.0, 3.0], [3.0, 2.0, 3.0]])
    V = np.concatenate([base, r.uniform(3 * 600, -5, 5).reshape(600, 3)])
    F = np.concatenate([np.tile([[0, 1, 2]], (700, 1)), np.tile([[1, 3, 4]], (500, 1)),
                        5 + np.arange(600).reshape(200, 3)]).astype(np.uint32)

    def fill(s):
        s.add_spheres(c, rad)
        s.add_mesh(V, F)
    print("device build: degenerate centers", flush=True)
    h, d = _pair(fill)
    assert h.build_stats(1)["bad_splits"] > 0 and h.build_stats(0)["bad_splits"] > 0
    _assert_same_build(h, d, [0, 1])


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 255, 256, 257, 258, 513, 4097])
def test_small_and_boundary_counts(dev, n):
    c, rad = W.sphere_field(n, seed=n)
    print(f"device build: {n} spheres", flush=True)
    _assert_same_build(*_pair(lambda s: s.add_spheres(c, rad)), [1])


def test_negative_zero_falls_back_to_host(dev):
    V, F = W.icosphere(3)
    V = V.copy()
    V[0, 0] = -0.0
    h, d = _pair(_mesh(V, F))
    _assert_same_build(h, d, [0], on_device=False)


def test_scene_primitives_example(dev):
    sp = W.scene_primitives()

    def fill(s):
        s.add_cylinders(*sp["cylinders"])
        s.add_spheres(*sp["spheres"])
    h, d = _pair(fill)
    _assert_same_build(h, d, [1, 3], on_device=None)


def test_hits_through_device_built_scene(dev, oracle_mod):
    m = W.mixed_scene(20_000, 5_000, 1_000, 100, seed=4)
    fill = _fill_mixed(m)
    from blazert_amd import Scene
    d = Scene()
    fill(d)
    d.commit(device=True)
    o = oracle_mod.Scene()
    fill(o)
    o.commit()
    org, dirs = W.mixed_rays(1 << 15, seed=5)
    O = torch.from_numpy(org).to(dev)
    D = torch.from_numpy(dirs).to(dev)
    g = d.intersect(O, D, normals=True)
    torch.cuda.synchronize()
    ref = o.intersect_batch(org, dirs, trace=True)
    ok = ref["trace"]["overflow"] == 0
    for k in ("t", "u", "v"):
        a, b = g[k].cpu().numpy()[ok], ref[k][ok]
        assert np.array_equal(a.view(np.uint64), b.view(np.uint64)), k
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(g[k].cpu().numpy().view(np.uint32)[ok], ref[k][ok]), k


def test_leaf_tuner_rebuilds_on_device(dev):
    from blazert_amd import Scene
    V, F = W.icosphere(5)
    d = Scene()
    d.add_mesh(V, F)
    d.commit(device=True)
    org, dirs = W.incoherent_rays(1 << 14, 3, 2.0)
    from blazert_amd.autotune import tune_leaf
    r = tune_leaf(d, torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev), leaf_sizes=(2, 3))
    leaf = r["min_leaf_primitives"]
    assert leaf in (2, 3)
    h = Scene(min_leaf_primitives=leaf)
    h.add_mesh(V, F)
    h.commit()
    _assert_same_build(h, d, [0])


def test_icosphere8_full(dev):
    """The 1,310,720-triangle benchmark scene: byte-identical arrays; commit times printed."""
    from blazert_amd import Scene
    V, F = W.icosphere(8)
    h = Scene()
    h.add_mesh(V, F)
    t0 = time.perf_counter()
    h.commit()
    th = time.perf_counter() - t0
    d = Scene()
    d.add_mesh(V, F)
    d.commit(device=True)  # warm-up of the build kernels
    d = Scene()
    d.add_mesh(V, F)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    d.commit(device=True)
    td = time.perf_counter() - t0
    print(f"commit 1.31M triangles: host {th * 1e3:.1f} ms, device {td * 1e3:.1f} ms", flush=True)
    _assert_same_build(h, d, [0])


NODE4 = np.dtype([("lo", np.float32, (3, 4)), ("hi", np.float32, (3, 4)), ("ref", np.uint32, 4),
                  ("cnt", np.uint32, 4)])
LEAF_BIT = 0x80000000
EMPTY_SLOT = 0xFFFFFFFF


def _f32_down(x):
    f = np.float32(x)
    return np.nextafter(f, np.float32(-np.inf)) if float(f) > x else f


def _f32_up(x):
    f = np.float32(x)
    return np.nextafter(f, np.float32(np.inf)) if float(f) < x else f


def _expected_node4(T):
    """Records of the 4-wide split-precision tree as documented: from a node's two children, the internal slot of
    largest surface area (the first of equal ones) is replaced in place by its two children until four slots are
    filled; records are numbered depth first, each record's internal slots consecutively when it is written."""
    def area(i):
        d = T["max"][i] - T["min"][i]
        return d[0] * d[1] + d[1] * d[2] + d[2] * d[0]

    def slots(b):
        sl = [int(T["children"][b][0]), int(T["children"][b][1])]
        while len(sl) < 4:
            cand = [(c, area(x)) for c, x in enumerate(sl) if not T["leaf"][x]]
            if not cand:
                break
            best = cand[0]
            for c, a in cand[1:]:
                if a > best[1]:
                    best = (c, a)
            e = sl[best[0]]
            sl[best[0]:best[0] + 1] = [int(T["children"][e][0]), int(T["children"][e][1])]
        return sl

    recs, work = [None], [(0, 0)]
    while work:
        b, k = work.pop()
        sl, refs, cnts, inner = slots(b), [], [], []
        for x in sl:
            if T["leaf"][x]:
                refs.append(LEAF_BIT | int(T["prim_begin"][x]))
                cnts.append(int(T["prim_count"][x]))
            else:
                refs.append(len(recs))
                cnts.append(0)
                recs.append(None)
                inner.append((x, refs[-1]))
        recs[k] = (sl, refs, cnts)
        work.extend(reversed(inner))
    return recs


@pytest.mark.parametrize("leaf", [1, 4])
def test_node4_records_follow_the_greedy_collapse(dev, leaf):
    """Both builders' 4-wide records against a restatement of the collapse rule: slots, references, counts, boxes
    rounded outward to FP32 and the empty-slot encoding, record for record; the record count is below the
    two-levels-at-a-time count (one record per internal node at even depth)."""
    V, F = W.icosphere(5)
    h, d = _pair(_mesh(V, F), min_leaf_primitives=leaf)
    T, _ = h.export_tree(0)
    want = _expected_node4(T)
    depth = np.zeros(len(T), dtype=np.int64)
    for i in range(len(T)):
        if not T["leaf"][i]:
            depth[T["children"][i]] = depth[i] + 1
    assert len(want) < int(np.sum((T["leaf"] == 0) & (depth % 2 == 0)))
    for s in (h, d):
        got = s.read_device_array(0, capi.ARRAY_NODES4).view(NODE4)
        assert len(got) == len(want)
        for k, (sl, refs, cnts) in enumerate(want):
            r = got[k]
            assert list(r["ref"][:len(sl)]) == refs and list(r["cnt"][:len(sl)]) == cnts, k
            for c, x in enumerate(sl):
                for a in range(3):
                    assert r["lo"][a][c] == _f32_down(T["min"][x][a]) and r["hi"][a][c] == _f32_up(T["max"][x][a])
            for c in range(len(sl), 4):
                assert r["cnt"][c] == EMPTY_SLOT and r["ref"][c] == 0
                assert np.all(r["lo"][:, c] == np.inf) and np.all(r["hi"][:, c] == -np.inf)
