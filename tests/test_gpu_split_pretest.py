"""The split path's FP32 triangle pre-test (tri_filter in blazert_amd/csrc/kernels.hip) only skips FP64 triangle
tests that would reject: with and without the pre-test records (BLAZERT_AMD_SPLIT_PRETEST=0 at upload) the
split-precision hits are identical bit for bit, on every ray. The rays are made to sit on the pre-test's error
bound: aimed at vertices and edge points (u, v or w within rounding of 0), grazing the triangles' planes (det near
0), with tmin / tmax windows ending at the hit, at scene scales from 1e-6 to 1e11."""
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
