"""The split path's FP32 triangle and sphere pre-tests (tri_filter and sph_filter in blazert_amd/csrc/kernels.hip)
only skip FP64 tests that would reject: with and without the pre-test records (BLAZERT_AMD_SPLIT_PRETEST=0 at upload)
the split-precision hits are identical bit for bit, on every ray. The rays are made to sit on the pre-tests' error
bounds: aimed at triangle vertices and edge points (u, v or w within rounding of 0), grazing the triangles' planes
(det near 0), grazing spheres within rounding of their radius, with tmin / tmax windows ending at the hit, at scene
scales from 1e-6 to 1e11 and far from the origin."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import Scene, workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")


def scene(v, f, pretest):
    old = os.environ.get("BLAZERT_AMD_SPLIT_PRETEST")
    os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = "1" if pretest else "0"
    try:
        s = Scene()
        s.add_mesh(v, f)
        s.commit()
        s.upload()
    finally:
        if old is None:
            del os.environ["BLAZERT_AMD_SPLIT_PRETEST"]
        else:
            os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = old
    return s


def trace(s, org, dirs, dev, tmin=None, tmax=None):
    col = lambda a: torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    h = s.intersect(tuple(col(org[:, k]) for k in range(3)), tuple(col(dirs[:, k]) for k in range(3)),
                    None if tmin is None else col(tmin), None if tmax is None else col(tmax), precision=1,
                    normals=True)
    torch.cuda.synchronize()
    return {k: h[k].cpu().numpy() for k in h}


def adversarial_rays(v, f, n, seed):
    """Rays from random origins through points of the triangles where the barycentric test is decided by rounding:
    vertices, edge points, and points just inside / outside an edge; plus rays nearly in a triangle's plane."""
    r = np.random.default_rng(seed)
    tri = f[r.integers(len(f), size=n)]
    p0, p1, p2 = v[tri[:, 0]], v[tri[:, 1]], v[tri[:, 2]]
    kind = r.integers(4, size=n)
    a = r.random(n)
    eps = np.where(r.random(n) < 0.5, -1.0, 1.0) * 10.0 ** r.uniform(-15, -6, n)
    # barycentric targets: a vertex, an edge point, just off an edge, a random interior point
    b1 = np.select([kind == 0, kind == 1, kind == 2], [np.zeros(n), a, a], r.random(n) * 0.5)
    b2 = np.select([kind == 0, kind == 1, kind == 2], [np.zeros(n), 1.0 - a, 1.0 - a + eps], r.random(n) * 0.5)
    target = p0 + b1[:, None] * (p1 - p0) + b2[:, None] * (p2 - p0)
    ext = np.abs(v).max()
    org = target + r.normal(size=(n, 3)) * ext * 0.5
    dirs = target - org
    # a quarter of the rays nearly in the plane of their triangle
    nrm = np.cross(p1 - p0, p2 - p0)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    graze = r.random(n) < 0.25
    inplane = dirs - (dirs * nrm).sum(1, keepdims=True) * nrm
    dirs = np.where(graze[:, None], inplane + nrm * np.linalg.norm(inplane, axis=1, keepdims=True) *
                    10.0 ** r.uniform(-12, -3, n)[:, None], dirs)
    return org, dirs


@pytest.mark.parametrize("scale", [1e-6, 1.0, 3e3, 1e11])
def test_pretest_never_changes_a_hit(dev, scale):
    v, f = W.icosphere(6, seed=3)
    v = v * scale
    with_pt, without = scene(v, f, True), scene(v, f, False)
    n = 1 << 18
    org, dirs = adversarial_rays(v, f, n, seed=int(np.log10(scale) + 20))
    inc_o, inc_d = W.incoherent_rays(n, 5, 2.0)
    org = np.concatenate([org, inc_o * scale])
    dirs = np.concatenate([dirs, inc_d])
    a, b = trace(with_pt, org, dirs, dev), trace(without, org, dirs, dev)
    for k in ("t", "u", "v", "normal"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["prim_id"] != -1).sum() > n // 2
    # windows that end at the hit: tmax just above / at / below the hit distance, tmin at it
    t = b["t"]
    hit = b["prim_id"] != -1
    m = np.nonzero(hit)[0][:n // 2]
    o2, d2, th = org[m], dirs[m], t[m]
    tmax = th * np.array([1 + 1e-15, 1.0, 1 - 1e-15])[np.arange(len(m)) % 3]
    tmin = np.where(np.arange(len(m)) % 2 == 0, 0.0, th * (1 - 1e-14))
    a2 = trace(with_pt, o2, d2, dev, tmin=tmin, tmax=tmax)
    b2 = trace(without, o2, d2, dev, tmin=tmin, tmax=tmax)
    for k in ("t", "prim_id"):
        assert np.array_equal(a2[k].view(np.uint64 if k == "t" else np.uint32),
                              b2[k].view(np.uint64 if k == "t" else np.uint32)), k
    with_pt.close()
    without.close()


def sphere_scene(c, r, pretest):
    old = os.environ.get("BLAZERT_AMD_SPLIT_PRETEST")
    os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = "1" if pretest else "0"
    try:
        s = Scene()
        s.add_spheres(c, r)
        s.commit()
        s.upload()
    finally:
        if old is None:
            del os.environ["BLAZERT_AMD_SPLIT_PRETEST"]
        else:
            os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = old
    return s


def tangent_rays(c, r, n, seed, scale):
    """Rays whose line passes the center of a random sphere at a distance within rounding of its radius (grazing
    the sphere from either side), rays from inside a sphere, and rays aimed at a sphere's center."""
    g = np.random.default_rng(seed)
    k = g.integers(len(c), size=n)
    d = g.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    side = g.normal(size=(n, 3))
    side -= (side * d).sum(1, keepdims=True) * d
    side /= np.linalg.norm(side, axis=1, keepdims=True)
    kind = g.integers(3, size=n)
    rel = 1.0 + np.where(g.random(n) < 0.5, -1.0, 1.0) * 10.0 ** g.uniform(-16, -5, n)
    off = np.select([kind == 0, kind == 1], [r[k] * rel, r[k] * g.random(n) * 0.9], np.zeros(n))
    through = c[k] + side * off[:, None]
    back = np.where(kind == 1, 0.0, g.uniform(0.5, 20.0, n) * scale)  # kind 1 starts inside the sphere
    return through - d * back[:, None], d


@pytest.mark.parametrize("scale,shift", [(1.0, 0.0), (1e-5, 0.0), (1.0, 3e3), (1e3, 1e8)])
def test_sphere_pretest_never_changes_a_hit(dev, scale, shift):
    """The split path's FP32 sphere pre-test (sph_filter) only skips FP64 sphere tests that report no hit: split
    hits with and without the pre-test records are identical bit for bit, on rays that graze spheres within rounding
    of their radius, start inside them or pass their centers, at several scales and far from the origin."""
    g = np.random.default_rng(int(np.log10(scale) + 30) + int(shift > 0))
    m = 20000
    c = (g.uniform(-10, 10, size=(m, 3)) * scale + shift).astype(np.float64)
    r = (10.0 ** g.uniform(-3, 0, m) * scale).astype(np.float64)
    with_pt, without = sphere_scene(c, r, True), sphere_scene(c, r, False)
    n = 1 << 17
    org, dirs = tangent_rays(c, r, n, 7, scale)
    inc_o, inc_d = W.incoherent_rays(n, 9, 12.0)
    org = np.concatenate([org, inc_o * scale + shift])
    dirs = np.concatenate([dirs, inc_d])
    a, b = trace(with_pt, org, dirs, dev), trace(without, org, dirs, dev)
    for k in ("t", "u", "v", "normal"):
        assert np.array_equal(a[k].view(np.uint64), b[k].view(np.uint64)), k
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(a[k], b[k]), k
    assert (a["prim_id"] != -1).sum() > n // 2
    # windows that end at the hit
    t = b["t"]
    hit = np.nonzero(b["prim_id"] != -1)[0][:n // 2]
    th = t[hit]
    tmax = th * np.array([1 + 1e-15, 1.0, 1 - 1e-15])[np.arange(len(hit)) % 3]
    tmin = np.where(np.arange(len(hit)) % 2 == 0, 0.0, th * (1 - 1e-14))
    a2 = trace(with_pt, org[hit], dirs[hit], dev, tmin=tmin, tmax=tmax)
    b2 = trace(without, org[hit], dirs[hit], dev, tmin=tmin, tmax=tmax)
    for k in ("t", "prim_id"):
        assert np.array_equal(a2[k].view(np.uint64 if k == "t" else np.uint32),
                              b2[k].view(np.uint64 if k == "t" else np.uint32)), k
    with_pt.close()
    without.close()


def cylinder_scene(c, a, b, h, rot, pretest):
    old = os.environ.get("BLAZERT_AMD_SPLIT_PRETEST")
    os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = "1" if pretest else "0"
    try:
        s = Scene()
        s.add_cylinders(c, a, b, h, rot)
        s.commit()
        s.upload()
    finally:
        if old is None:
            del os.environ["BLAZERT_AMD_SPLIT_PRETEST"]
        else:
            os.environ["BLAZERT_AMD_SPLIT_PRETEST"] = old
    return s


def cylinder_rays(c, a, b, h, rot, n, seed, scale):
    """Rays that graze a random cylinder's shell (through a surface point, along a tangent of its ellipse), run
    parallel to its axis through a point of the rim, cross a cap's rim, or start inside it. Local points are mapped
    to the world as c + R p, the frame of blazert's cylinder test (cylinder.h:181-495)."""
    g = np.random.default_rng(seed)
    k = g.integers(len(c), size=n)
    th = g.uniform(0, 2 * np.pi, n)
    z = g.uniform(-0.5, 0.5, n) * h[k]
    rim = np.stack([a[k] * np.cos(th), b[k] * np.sin(th), z], 1)
    tangent = np.stack([-a[k] * np.sin(th), b[k] * np.cos(th), g.normal(size=n) * 1e-3], 1)
    kind = g.integers(4, size=n)
    p = rim.copy()
    p[kind == 2, 2] = np.where(g.random((kind == 2).sum()) < 0.5, 0.5, -0.5) * h[k][kind == 2]
    p[kind == 3] *= g.uniform(0, 0.9, ((kind == 3).sum(), 1))
    d = np.select([kind[:, None] == 0, kind[:, None] == 1], [tangent, np.array([0.0, 0.0, 1.0])],
                  g.normal(size=(n, 3)))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    R = rot.reshape(-1, 3, 3)[k]
    pw = c[k] + np.einsum("nij,nj->ni", R, p)
    dw = np.einsum("nij,nj->ni", R, d)
    dw /= np.linalg.norm(dw, axis=1, keepdims=True)
    back = np.where(kind == 3, 0.0, g.uniform(0.5, 20.0, n) * scale)
    return pw - dw * back[:, None], dw


@pytest.mark.parametrize("scale,shift", [(1.0, 0.0), (1e-5, 0.0), (1e3, 1e8)])
def test_cylinder_pretest_never_changes_a_hit(dev, scale, shift):
    """The split path's bounding-sphere pre-test of cylinders only skips FP64 cylinder tests that report no hit: split
    hits with and without the pre-test records are identical bit for bit, on rays grazing the shell, parallel to the
    axis through the rim, crossing cap rims and starting inside, for orthonormal rotations and for rotations that
    are not (whose cylinders the pre-test never rejects)."""
    g = np.random.default_rng(int(np.log10(scale) + 40) + int(shift > 0))
    m = 5000
    c = (g.uniform(-10, 10, size=(m, 3)) * scale + shift).astype(np.float64)
    a = 10.0 ** g.uniform(-2, 0, m) * scale
    b = 10.0 ** g.uniform(-2, 0, m) * scale
    h = 10.0 ** g.uniform(-1, 0.5, m) * scale
    q, _ = np.linalg.qr(g.normal(size=(m, 3, 3)))
    q[m // 10:m // 5] *= 1.01  # not orthonormal
    rot = np.ascontiguousarray(q.reshape(m, 9))
    with_pt, without = cylinder_scene(c, a, b, h, rot, True), cylinder_scene(c, a, b, h, rot, False)
    n = 1 << 17
    org, dirs = cylinder_rays(c, a, b, h, rot, n, 11, scale)
    inc_o, inc_d = W.incoherent_rays(n, 13, 12.0)
    org = np.concatenate([org, inc_o * scale + shift])
    dirs = np.concatenate([dirs, inc_d])
    x, y = trace(with_pt, org, dirs, dev), trace(without, org, dirs, dev)
    for k in ("t", "u", "v", "normal"):
        assert np.array_equal(x[k].view(np.uint64), y[k].view(np.uint64)), k
    for k in ("prim_id", "geom_id"):
        assert np.array_equal(x[k], y[k]), k
    assert (x["prim_id"] != -1).sum() > n // 4
    with_pt.close()
    without.close()
