"""User-defined collections on the GPU: device intersectors plugged into the header-only traversal of
include/blazert_amd/user_geometry.h (the example plugin examples/user_geometry/user_geometry.hip, built by
__graft_entry__.build()).

* UserSpheres: the reference's sphere re-expressed as a user collection over a brt_bvh tree must give the engine's
  built-in sphere path (itself bit-identical to the oracle) bit for bit: t, u, v, prim_id, geom_id and normal, for
  closest hit, any hit and tmin / tmax windows, in T = double and T = float, and with the tree built through the
  header's template constructor from a host collection object.
* UserBoxes: a primitive the engine does not have (boxes hit at their entry distance) against a brute-force minimum
  over every box in numpy, with the same arithmetic.
"""
import ctypes as C
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import BVH, Scene, capi  # noqa: E402
from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLUGIN = os.path.join(ROOT, "examples", "user_geometry", "libuser_geometry.so")
INVALID = 0xFFFFFFFF


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


@pytest.fixture(scope="module")
def plugin():
    L = C.CDLL(PLUGIN)
    P, U, U64 = C.c_void_p, C.c_uint, C.c_uint64
    L.ug_spheres_build.argtypes = [P, P, U, C.POINTER(capi.BuildOptions), C.POINTER(C.c_void_p)]
    L.ug_spheres_intersect.argtypes = [P, P, P, U, C.POINTER(capi.Rays), C.POINTER(capi.Hits), U64, U, P]
    L.ug_spheres_intersect_f32.argtypes = [P, P, P, U, C.POINTER(capi.Rays), C.POINTER(capi.Hits), U64, U, P]
    L.ug_boxes_intersect.argtypes = [P, P, P, U, C.POINTER(capi.Rays), C.POINTER(capi.Hits), U64, P]
    return L


def _rays(org, dirs, dev, tdt, tmin=None, tmax=None):
    o = torch.from_numpy(np.ascontiguousarray(org.T)).to(dev, tdt).contiguous()
    d = torch.from_numpy(np.ascontiguousarray(dirs.T)).to(dev, tdt).contiguous()
    tmn = None if tmin is None else torch.from_numpy(np.ascontiguousarray(tmin)).to(dev, tdt)
    tmx = None if tmax is None else torch.from_numpy(np.ascontiguousarray(tmax)).to(dev, tdt)
    R = capi.Rays(*[o[k].data_ptr() for k in range(3)], *[d[k].data_ptr() for k in range(3)],
                  tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
    return R, (o, d, tmn, tmx)


def _hits(n, dev, tdt):
    h = {k: torch.full((n,), 7.0, dtype=tdt, device=dev) for k in ("t", "u", "v")}
    h["prim_id"] = torch.full((n,), 5, dtype=torch.int32, device=dev)
    h["geom_id"] = torch.full((n,), 5, dtype=torch.int32, device=dev)
    h["normal"] = torch.full((3, n), 7.0, dtype=tdt, device=dev)
    H = capi.Hits(h["t"].data_ptr(), h["u"].data_ptr(), h["v"].data_ptr(), h["prim_id"].data_ptr(),
                  h["geom_id"].data_ptr(), *[h["normal"][k].data_ptr() for k in range(3)])
    return h, H


def _np(h):
    out = {k: h[k].cpu().numpy() for k in ("t", "u", "v")}
    out["prim_id"] = h["prim_id"].cpu().numpy().view(np.uint32)
    out["geom_id"] = h["geom_id"].cpu().numpy().view(np.uint32)
    out["normal"] = h["normal"].cpu().numpy()
    return out


def _assert_same(a, b):
    for k in ("t", "u", "v", "normal"):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        ui = np.uint64 if x.dtype == np.float64 else np.uint32
        same = (x.view(ui) == y.view(ui)) | (np.isnan(x) & np.isnan(y))
        assert same.all(), f"{k}: {int((~same).sum())} mismatches"
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(a[k], b[k]), k


def _stream():
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def _plugin_spheres(plugin, handle, c_dev, r_dev, n, org, dirs, dev, tdt, any_hit=False, tmin=None, tmax=None):
    R, keep = _rays(org, dirs, dev, tdt, tmin, tmax)
    h, H = _hits(len(org), dev, tdt)
    fn = plugin.ug_spheres_intersect if tdt == torch.float64 else plugin.ug_spheres_intersect_f32
    rc = fn(handle, c_dev.data_ptr(), r_dev.data_ptr(), n, C.byref(R), C.byref(H), len(org), int(any_hit), _stream())
    assert rc == 0, capi.lib().brt_last_error()
    torch.cuda.synchronize()
    return _np(h)


def _builtin(scene, org, dirs, dev, tdt, **kw):
    O = torch.from_numpy(np.ascontiguousarray(org)).to(dev, tdt)
    D = torch.from_numpy(np.ascontiguousarray(dirs)).to(dev, tdt)
    tm = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev, tdt) for k, v in kw.items() if k in ("tmin", "tmax")
          and v is not None}
    h = scene.intersect(O, D, normals=True, any_hit=kw.get("any_hit", False), **tm)
    torch.cuda.synchronize()
    return _np(h)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_user_spheres_equal_builtin_spheres(dev, plugin, dtype):
    tdt = torch.float64 if dtype == np.float64 else torch.float32
    c, r = W.sphere_field(20_000, seed=31)
    c, r = c.astype(dtype), r.astype(dtype)
    s = Scene(dtype=dtype)
    s.add_spheres(c, r)
    s.commit()
    bvh = BVH(box_min=c - r[:, None], box_max=c + r[:, None], centers=c, dtype=dtype)
    c_dev = torch.from_numpy(np.ascontiguousarray(c)).to(dev)
    r_dev = torch.from_numpy(np.ascontiguousarray(r)).to(dev)
    n = len(r)
    org, dirs = W.incoherent_rays(1 << 15, 32, 12.0)
    print(f"user spheres ({np.dtype(dtype).name}): {n} spheres, {len(org)} incoherent rays", flush=True)
    got = _plugin_spheres(plugin, bvh.handle, c_dev, r_dev, n, org, dirs, dev, tdt)
    ref = _builtin(s, org, dirs, dev, tdt)
    assert (ref["prim_id"] != INVALID).sum() > 1000
    _assert_same(got, ref)
    # any hit
    got = _plugin_spheres(plugin, bvh.handle, c_dev, r_dev, n, org, dirs, dev, tdt, any_hit=True)
    _assert_same(got, _builtin(s, org, dirs, dev, tdt, any_hit=True))
    # tmin / tmax windows
    tmin = np.where(W.Rng(33).uniform(len(org)) < 0.3, 2.0, 0.0).astype(dtype)
    tmax = np.where(W.Rng(34).uniform(len(org)) < 0.3, 8.0, np.finfo(dtype).max).astype(dtype)
    got = _plugin_spheres(plugin, bvh.handle, c_dev, r_dev, n, org, dirs, dev, tdt, tmin=tmin, tmax=tmax)
    _assert_same(got, _builtin(s, org, dirs, dev, tdt, tmin=tmin, tmax=tmax))
    # a camera (coherent) batch
    corg, cdirs = W.camera_rays(200, 150, (0.0, 0.0, 30.0))
    got = _plugin_spheres(plugin, bvh.handle, c_dev, r_dev, n, corg, cdirs, dev, tdt)
    _assert_same(got, _builtin(s, corg, cdirs, dev, tdt))


def test_template_builder_on_gpu(dev, plugin):
    """BVH<double> built by the header's template constructor from a host collection object."""
    c, r = W.sphere_field(5_000, seed=35)
    c, r = np.ascontiguousarray(c), np.ascontiguousarray(r)
    h = C.c_void_p()
    opts = capi.BuildOptions(2, 28, 32)
    assert plugin.ug_spheres_build(c.ctypes.data, r.ctypes.data, len(r), C.byref(opts), C.byref(h)) == 0
    try:
        s = Scene(min_leaf_primitives=2, bin_size=32)
        s.add_spheres(c, r)
        s.commit()
        org, dirs = W.incoherent_rays(1 << 14, 36, 12.0)
        got = _plugin_spheres(plugin, h, torch.from_numpy(c).to(dev), torch.from_numpy(r).to(dev), len(r), org, dirs,
                              dev, torch.float64)
        _assert_same(got, _builtin(s, org, dirs, dev, torch.float64))
    finally:
        capi.lib().brt_bvh_destroy(h)


def _boxes(n, seed):
    rng = W.Rng(seed)
    ctr = rng.uniform(3 * n, -20.0, 20.0).reshape(n, 3)
    half = np.exp(rng.uniform(3 * n, -3.0, 0.5)).reshape(n, 3)
    return ctr - half, ctr + half, ctr


def _bruteforce_boxes(lo, hi, org, dirs):
    """The plugin's box test in numpy for every (ray, box): entry distance, the strict first-axis-wins choice of the
    entry axis, and the closest hit over all boxes (the lowest index on equal distances)."""
    n = len(org)
    inv = 1.0 / dirs
    best = np.full(n, np.finfo(np.float64).max)
    prim = np.full(n, INVALID, np.uint32)
    axis = np.zeros(n, np.int64)
    for s in range(0, len(lo), 256):
        L, H = lo[s:s + 256], hi[s:s + 256]
        a = (L[None, :, :] - org[:, None, :]) * inv[:, None, :]
        c = (H[None, :, :] - org[:, None, :]) * inv[:, None, :]
        mn, mx = np.where(a < c, a, c), np.where(a < c, c, a)
        tn = np.full(a.shape[:2], -np.inf)
        ax = np.zeros(a.shape[:2], np.int64)
        tf = np.full(a.shape[:2], np.inf)
        for k in range(3):
            up = mn[:, :, k] > tn
            tn = np.where(up, mn[:, :, k], tn)
            ax = np.where(up, k, ax)
            tf = np.where(mx[:, :, k] < tf, mx[:, :, k], tf)
        ok = (tn <= tf) & (tn > 0.0)
        tn = np.where(ok, tn, np.inf)
        j = np.argmin(tn, axis=1)
        tj = tn[np.arange(n), j]
        better = tj < best
        best = np.where(better, tj, best)
        prim = np.where(better, (s + j).astype(np.uint32), prim)
        axis = np.where(better, ax[np.arange(n), j], axis)
    hit = prim != INVALID
    normal = np.zeros((3, n))
    sgn = np.where(inv[np.arange(n), axis] < 0.0, 1.0, -1.0)
    normal[axis[hit], np.nonzero(hit)[0]] = sgn[hit]
    return {"t": best, "u": np.zeros(n), "v": np.zeros(n), "prim_id": prim,
            "geom_id": np.where(hit, 7, INVALID).astype(np.uint32), "normal": normal}


def test_user_boxes_bruteforce(dev, plugin):
    lo, hi, ctr = _boxes(3000, 37)
    bvh = BVH(box_min=lo, box_max=hi, centers=ctr)
    rng = W.Rng(38)
    m = 4096
    org = rng.unit_vectors(m) * 45.0
    target = rng.uniform(3 * m, -15.0, 15.0).reshape(m, 3)
    dirs = target - org  # no zero components: every 1 / d is finite
    assert (dirs != 0).all()
    R, keep = _rays(org, dirs, dev, torch.float64)
    h, H = _hits(m, dev, torch.float64)
    lo_d = torch.from_numpy(np.ascontiguousarray(lo)).to(dev)
    hi_d = torch.from_numpy(np.ascontiguousarray(hi)).to(dev)
    print(f"user boxes: {len(lo)} boxes, {m} rays", flush=True)
    rc = plugin.ug_boxes_intersect(bvh.handle, lo_d.data_ptr(), hi_d.data_ptr(), len(lo), C.byref(R), C.byref(H), m,
                                   _stream())
    assert rc == 0, capi.lib().brt_last_error()
    torch.cuda.synchronize()
    got = _np(h)
    ref = _bruteforce_boxes(lo, hi, org, dirs)
    assert (ref["prim_id"] != INVALID).sum() > 1000
    _assert_same(got, ref)


def test_irregular_boxes_flag(dev):
    """brt_bvh_view.irregular_boxes marks trees with an inverted (min > max) or NaN node box, which the device
    traversal then tests with the sign-selected general slab form. A node box is the union of its primitives' boxes,
    so one primitive per leaf makes a primitive's box a node box."""
    lo, hi, ctr = _boxes(100, 44)
    assert BVH(box_min=lo, box_max=hi, centers=ctr, min_leaf_primitives=1).device_view().irregular_boxes == 0
    lo2, hi2 = lo.copy(), hi.copy()
    lo2[7], hi2[7] = hi[7], lo[7]
    assert BVH(box_min=lo2, box_max=hi2, centers=ctr, min_leaf_primitives=1).device_view().irregular_boxes == 1
    lo3 = lo.copy()
    lo3[3, 1] = np.nan
    assert BVH(box_min=lo3, box_max=hi, centers=ctr, min_leaf_primitives=1).device_view().irregular_boxes == 1


def test_empty_collection(dev, plugin):
    z = np.zeros((0, 3))
    bvh = BVH(box_min=z, box_max=z, centers=z)
    v = bvh.device_view()
    assert v.num_prims == 0 and v.nodes is None
    org, dirs = W.incoherent_rays(1000, 39, 5.0)
    e = torch.zeros(1, dtype=torch.float64, device=dev)
    got = _plugin_spheres(plugin, bvh.handle, e, e, 0, org, dirs, dev, torch.float64)
    assert (got["t"] == np.finfo(np.float64).max).all()
    assert (got["prim_id"] == INVALID).all() and (got["geom_id"] == INVALID).all()
    assert (got["u"] == 0).all() and (got["normal"] == 0).all()
