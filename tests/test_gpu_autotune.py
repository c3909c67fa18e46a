"""Autotuner (brt_autotune_launch / brt_batch_autotune / brt_scene_tune_leaf): a tuned configuration gives the same
hits as the defaults, and a leaf-size-tuned scene matches the oracle built with the chosen leaf size."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401
    return torch.device("cuda:0")


@pytest.fixture(autouse=True)
def tmp_cache(tmp_path, monkeypatch):
    monkeypatch.setenv("BLAZERT_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))


def _hits(h):
    return {"t": h["t"].cpu().numpy(), "prim_id": h["prim_id"].cpu().numpy().view(np.uint32),
            "u": h["u"].cpu().numpy(), "v": h["v"].cpu().numpy()}


def _same(a, b):
    return all(np.array_equal(np.ascontiguousarray(a[k]).view(np.uint64 if a[k].dtype == np.float64 else np.uint32),
                              np.ascontiguousarray(b[k]).view(np.uint64 if b[k].dtype == np.float64 else np.uint32))
               for k in a)


@pytest.mark.parametrize("precision", [0, 1])
def test_tuned_launch_same_hits(dev, precision):
    from blazert_amd import Scene
    from blazert_amd.autotune import LAUNCH_KEYS, tune_launch
    v, f = W.icosphere(5)
    s = Scene()
    s.add_mesh(v, f)
    s.commit()
    org, dirs = W.icosphere_incoherent_rays(1 << 16, seed=3)
    O, D = torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)
    res = tune_launch(s, O, D, precision=precision, repeats=2)
    print("tuned", res, flush=True)
    assert res["candidates"] == 1 + 4 * 4 * 3 * 2 + 2 and res["best_ms"] <= res["default_ms"]  # + pool sizes 32, 128
    base = _hits(s.intersect(O, D, precision=precision))
    tuned = _hits(s.intersect(O, D, precision=precision, **{k: res[k] for k in LAUNCH_KEYS}))
    assert _same(base, tuned)
    again = tune_launch(s, O, D, precision=precision)  # served from the cache
    assert again == res


def test_tuned_batch_same_hits(dev):
    from blazert_amd import Batch, Scene
    from blazert_amd.autotune import tune_batch
    items, ref = [], []
    for i, (lvl, n) in enumerate([(3, 5000), (5, 40000), (4, 777), (6, 20000)]):
        v, f = W.icosphere(lvl, seed=i)
        s = Scene()
        s.add_mesh(v, f)
        s.commit()
        org, dirs = W.incoherent_rays(n, seed=50 + i, box=2.0)
        O, D = torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)
        items.append((s, O, D))
        ref.append(_hits(s.intersect(O, D, precision=1)))
    b = Batch(items, precision=1)
    res = tune_batch(b, repeats=2)
    print("tuned batch", res, flush=True)
    assert res["candidates"] == 1 + 4 * 4 * 3 + 2
    stream = torch.cuda.Stream()
    b.capture(stream)
    for _ in range(2):
        for h in b.hits:
            h["t"].fill_(-1.0)
        torch.cuda.synchronize()
        b.replay(stream)
        stream.synchronize()
        for h, r in zip(b.hits, ref):
            assert _same(_hits(h), r)


def test_tune_leaf_matches_oracle(dev, oracle_mod):
    from blazert_amd import Scene
    from blazert_amd.autotune import tune_leaf
    v, f = W.icosphere(5, seed=9)
    s = Scene()
    s.add_mesh(v, f)
    org, dirs = W.icosphere_incoherent_rays(1 << 15, seed=4)
    O, D = torch.from_numpy(org).to(dev), torch.from_numpy(dirs).to(dev)
    res = tune_leaf(s, O, D, leaf_sizes=(1, 2, 4, 8), repeats=2)
    print("tuned leaf", res, flush=True)
    leaf = res["min_leaf_primitives"]
    assert leaf in (1, 2, 4, 8) and res["candidates"] == 4
    assert s.build_stats(0)["max_num_prims_per_leaf"] <= max(leaf, 1) or s.build_stats(0)["max_tree_depth"] >= 28
    o = oracle_mod.Scene(leaf, 28, 64)
    o.add_mesh(v, f)
    o.commit()
    g = _hits(s.intersect(O, D))
    r = o.intersect_batch(org, dirs, normals=False, trace=True)
    ok = r["trace"]["overflow"] == 0
    for k in ("t", "u", "v", "prim_id"):
        a, b = g[k][ok], r[k][ok]
        assert np.array_equal(a.view(np.uint64) if a.dtype == np.float64 else a,
                              b.view(np.uint64) if b.dtype == np.float64 else b), k
