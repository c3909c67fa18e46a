"""brt_intersect_noalloc: the allocation-free variant of brt_intersect, through the C ABI. Its launch counters and
ray-sort workspace live in the library's static device scratch. Its hits must equal brt_intersect's bit for bit
(every field), on two calls in a row, the second with fewer rays, for the FP64 path, the split-precision path (with
rays it defers to the FP64 path) and a sorted launch; calls it cannot serve are refused before any launch."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import Scene, capi, workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def scene(dev):
    v, f = W.icosphere(5)
    s = Scene()
    s.add_mesh(v * 2.0, f)
    c, r = W.sphere_field(512, seed=3)
    s.add_spheres(c * 0.25, r * 0.3)
    s.add_planes(np.array([[0.0, 0.0, -2.5]]), [3.0], [3.0], np.eye(3).reshape(1, 9))
    s.add_cylinders(np.array([[1.0, 0.5, 0.0], [-1.0, 0.0, 0.5]]), [0.3, 0.2], [0.3, 0.2], [2.0, 1.0],
                    np.stack([np.eye(3).reshape(9), W.rotation_to(np.array([[1.0, 1.0, 0.0]]))[0]]))
    s.commit()
    s.upload()
    return s


def rays(n, seed, dev):
    org, dirs = W.incoherent_rays(n, seed, 3.0)
    # a few hundred axis-parallel rays (zero direction components): the split path defers them to the FP64 path
    k = np.arange(0, n, 97)
    dirs[k, (k // 97) % 3] = 0.0
    return (tuple(torch.from_numpy(np.ascontiguousarray(org[:, j])).to(dev) for j in range(3)),
            tuple(torch.from_numpy(np.ascontiguousarray(dirs[:, j])).to(dev) for j in range(3)))


def call(fn, scene, O, D, dev, launch):
    n = O[0].numel()
    h = {k: torch.full((n,), -1.0, dtype=torch.float64, device=dev) for k in ("t", "u", "v", "nx", "ny", "nz")}
    h["prim_id"] = torch.full((n,), -3, dtype=torch.int32, device=dev)
    h["geom_id"] = torch.full((n,), -3, dtype=torch.int32, device=dev)
    R = capi.Rays(*[x.data_ptr() for x in O], *[x.data_ptr() for x in D], None, None)
    H = capi.Hits(*[h[k].data_ptr() for k in ("t", "u", "v", "prim_id", "geom_id", "nx", "ny", "nz")])
    torch.cuda.synchronize()
    rc = getattr(capi.lib(), fn)(scene.handle, C.byref(R), C.byref(H), n, C.byref(launch), None)
    torch.cuda.synchronize()
    return rc, h


def same(a, b):
    for k in ("t", "u", "v", "nx", "ny", "nz"):
        if not torch.equal(a[k].view(torch.int64), b[k].view(torch.int64)):
            return k
    for k in ("prim_id", "geom_id"):
        if not torch.equal(a[k], b[k]):
            return k
    return None


@pytest.mark.parametrize("precision,sort", [(0, 0), (1, 0), (0, 1), (1, 2)])
def test_noalloc_matches_intersect(dev, scene, precision, sort):
    L = capi.Launch(precision, 0, 0, capi.SCHEDULE_PERSISTENT, 0, sort, 0)
    for n, seed in ((300_000, 11), (70_001, 12)):  # two calls in a row, the second smaller
        O, D = rays(n, seed, dev)
        rc0, ref = call("brt_intersect", scene, O, D, dev, L)
        assert rc0 == capi.OK
        print(f"noalloc precision={precision} sort={sort} n={n}", flush=True)
        rc1, got = call("brt_intersect_noalloc", scene, O, D, dev, L)
        assert rc1 == capi.OK, capi.lib().brt_last_error()
        bad = same(got, ref)
        assert bad is None, f"field {bad} differs (n={n})"
        assert int((ref["prim_id"] != -1).sum()) > n // 10


def test_noalloc_refusals(dev, scene):
    lib = capi.lib()
    # a sorted launch beyond the scratch's sort capacity (2^20 rays): refused before any device work
    n = (1 << 20) + 1
    R = capi.Rays(1, 1, 1, 1, 1, 1, None, None)  # never dereferenced: the call is refused first
    H = capi.Hits(1, 1, 1, 1, 1, None, None, None)
    L = capi.Launch(0, 0, 0, capi.SCHEDULE_PERSISTENT, 0, capi.SORT_MORTON, 0)
    assert lib.brt_intersect_noalloc(scene.handle, C.byref(R), C.byref(H), n, C.byref(L), None) == \
        capi.ERR_UNSUPPORTED
    assert b"sorts at most" in lib.brt_last_error()
    # a committed scene that is not on the device yet: uploading would allocate
    s = Scene()
    s.add_spheres(np.zeros((1, 3)), [1.0])
    s.commit()
    L0 = capi.Launch(0, 0, 0, capi.SCHEDULE_PERSISTENT, 0, 0, 0)
    assert lib.brt_intersect_noalloc(s.handle, C.byref(R), C.byref(H), 10, C.byref(L0), None) == capi.ERR_STATE
    s.close()
