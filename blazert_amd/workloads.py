"""Synthetic scenes and ray sets for tests and benchmarks (numpy, float64, seeded).

One generator is shared by every consumer so CPU and GPU runs see identical inputs:
SplitMix64 over a counter -> uniform doubles from the top 53 bits; Gaussians by Box-Muller.

Workloads (sizes of the reference-path configurations):
* ``sphere_field``        1,024 spheres, centers U[-10,10]^3, radii U[0.1,0.5]; 512x512 pinhole rays from
                          (0,0,30) toward -z, 45 deg vertical FOV, tmin 0, tmax +inf.
* ``icosphere(8)``        subdivided icosahedron (1,310,720 triangles, 655,362 vertices), each vertex scaled
                          radially by (1 + 0.05 N(0,1)); coherent 1920x1080 camera rays from distance 4, and
                          4,194,304 incoherent rays (origins U[-2,2]^3, directions uniform on S^2).
* ``mixed_scene``         200k triangles (displaced icosphere patches), 50k spheres, 10k cylinders, 100 planes,
                          8,388,608 incoherent rays.
* ``scene_primitives``    the 2-cylinder / 2-sphere scene of blazert's examples/scene_primitives/main.cc:47-81.

The icosahedron and its subdivision order follow benchmarks/benchmark_helper/OriginSphere.h:11-119
(midpoint (v0+v1)*0.5 then normalize; children (f0,f3,f5), (f3,f1,f4), (f4,f2,f5), (f3,f4,f5)) with shared
edge midpoints de-duplicated, so every vertex exists once.
"""
from __future__ import annotations

import math

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    """n outputs of SplitMix64 started at state ``seed`` (counter based, vectorised)."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


class Rng:
    """Stream of uniform doubles / Gaussians from one SplitMix64 seed."""

    def __init__(self, seed: int):
        self.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
        self.pos = 0

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        z = splitmix64(self.seed, n, self.pos)
        self.pos += n
        u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        m = (n + 1) // 2
        u1 = self.uniform(m)
        u2 = self.uniform(m)
        r = np.sqrt(-2.0 * np.log1p(-u1))  # 1-u1 in (0,1]
        z = np.concatenate([r * np.cos(2 * np.pi * u2), r * np.sin(2 * np.pi * u2)])
        return z[:n]

    def unit_vectors(self, n: int) -> np.ndarray:
        """Directions uniform on S^2 (z uniform in [-1,1], azimuth uniform)."""
        z = self.uniform(n, -1.0, 1.0)
        phi = self.uniform(n, 0.0, 2 * np.pi)
        s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
        return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


# ------------------------------------------------------------------------------------------------ geometry
def icosahedron():
    t = float(np.float32((np.float32(1.0) + np.sqrt(np.float32(5.0))) / np.float32(2.0)))
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v, f


def subdivide(v: np.ndarray, f: np.ndarray):
    nv = len(v)
    e01 = np.sort(f[:, [0, 1]], axis=1)
    e12 = np.sort(f[:, [1, 2]], axis=1)
    e20 = np.sort(f[:, [2, 0]], axis=1)
    edges = np.concatenate([e01, e12, e20])
    key = edges[:, 0] * (nv + 1) + edges[:, 1]
    uniq, first, inv = np.unique(key, return_index=True, return_inverse=True)
    # midpoints numbered in order of first appearance (face order, edge order f3, f4, f5 per face)
    order = np.argsort(first, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    mid_id = nv + rank[inv]
    ue = edges[first[order]]
    mid = (v[ue[:, 0]] + v[ue[:, 1]]) * 0.5
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    nf = len(f)
    f3, f4, f5 = mid_id[:nf], mid_id[nf:2 * nf], mid_id[2 * nf:]
    f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
    nfaces = np.stack([np.stack([f0, f3, f5], 1), np.stack([f3, f1, f4], 1), np.stack([f4, f2, f5], 1),
                       np.stack([f3, f4, f5], 1)], axis=1).reshape(-1, 3)
    return np.concatenate([v, mid]), nfaces


def icosphere(level: int, seed: int = 7, displacement: float = 0.05):
    """Deduplicated icosphere; vertices scaled by (1 + displacement * N(0,1)). Returns (vertices, faces)."""
    v, f = icosahedron()
    for _ in range(level):
        v, f = subdivide(v, f)
    if displacement:
        g = Rng(seed).normal(len(v))
        v = v * (1.0 + displacement * g)[:, None]
    return np.ascontiguousarray(v), np.ascontiguousarray(f.astype(np.uint32))


def rotation_to(axis: np.ndarray) -> np.ndarray:
    """Row-major rotations whose third column is ``axis`` (local z -> axis); shape (n, 9)."""
    a = axis / np.linalg.norm(axis, axis=1, keepdims=True)
    helper = np.where(np.abs(a[:, :1]) < 0.9, np.array([[1.0, 0, 0]]), np.array([[0, 1.0, 0]]))
    x = np.cross(helper, a)
    x /= np.linalg.norm(x, axis=1, keepdims=True)
    y = np.cross(a, x)
    R = np.stack([x, y, a], axis=2)  # columns x, y, a
    return np.ascontiguousarray(R.reshape(-1, 9))


# ------------------------------------------------------------------------------------------------ rays
def camera_rays(width: int, height: int, eye, fov_deg: float = 45.0, look=(0.0, 0.0, -1.0)):
    """Pinhole rays through pixel centres; returns (origins, unit directions), pixel index = y*W + x."""
    look = np.asarray(look, float)
    look = look / np.linalg.norm(look)
    up = np.array([0.0, 1.0, 0.0]) if abs(look[1]) < 0.9 else np.array([0.0, 0.0, 1.0])
    right = np.cross(look, up)
    right /= np.linalg.norm(right)
    up = np.cross(right, look)
    th = math.tan(math.radians(fov_deg) / 2)
    aspect = width / height
    ys, xs = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    px = ((xs.reshape(-1) + 0.5) / width * 2 - 1) * th * aspect
    py = ((ys.reshape(-1) + 0.5) / height * 2 - 1) * th
    d = look[None] + px[:, None] * right[None] + py[:, None] * up[None]
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    o = np.broadcast_to(np.asarray(eye, float), d.shape).copy()
    return o, np.ascontiguousarray(d)


def incoherent_rays(n: int, seed: int, box: float):
    r = Rng(seed)
    o = r.uniform(3 * n, -box, box).reshape(3, n).T.copy()
    d = r.unit_vectors(n)
    return np.ascontiguousarray(o), np.ascontiguousarray(d)


# ------------------------------------------------------------------------------------------------ scenes
def sphere_field(n: int = 1024, seed: int = 1):
    r = Rng(seed)
    centers = r.uniform(3 * n, -10.0, 10.0).reshape(n, 3)
    radii = r.uniform(n, 0.1, 0.5)
    return np.ascontiguousarray(centers), radii


def sphere_field_rays(width: int = 512, height: int = 512):
    return camera_rays(width, height, eye=(0.0, 0.0, 30.0), fov_deg=45.0)


def icosphere_coherent_rays(width: int = 1920, height: int = 1080):
    return camera_rays(width, height, eye=(0.0, 0.0, 4.0), fov_deg=45.0)


def icosphere_incoherent_rays(n: int = 4 * 1024 * 1024, seed: int = 3):
    return incoherent_rays(n, seed, 2.0)


def mixed_scene(n_tri_target: int = 200_000, n_sph: int = 50_000, n_cyl: int = 10_000, n_pla: int = 100,
                seed: int = 11):
    """Mixed-primitive scene. Triangles: displaced icosphere patches (level 4, 5120 tris each) at offsets
    U[-20,20]^3 merged into one mesh (one add_mesh per scene is allowed, blazert/scene.h:230)."""
    r = Rng(seed)
    pv, pf = icosphere(4, seed=seed + 1)
    n_patch = max(1, int(round(n_tri_target / len(pf))))
    verts, faces = [], []
    scale = r.uniform(n_patch, 0.3, 1.5)
    off = r.uniform(3 * n_patch, -20.0, 20.0).reshape(n_patch, 3)
    for i in range(n_patch):
        verts.append(pv * scale[i] + off[i])
        faces.append(pf.astype(np.int64) + i * len(pv))
    V = np.ascontiguousarray(np.concatenate(verts))
    F = np.ascontiguousarray(np.concatenate(faces).astype(np.uint32))
    sc = r.uniform(3 * n_sph, -20.0, 20.0).reshape(n_sph, 3)
    sr = r.uniform(n_sph, 0.05, 0.3)
    cc = r.uniform(3 * n_cyl, -20.0, 20.0).reshape(n_cyl, 3)
    cr = r.uniform(n_cyl, 0.02, 0.2)
    ch = r.uniform(n_cyl, 0.5, 3.0)
    crot = rotation_to(r.unit_vectors(n_cyl))
    pc = r.uniform(3 * n_pla, -20.0, 20.0).reshape(n_pla, 3)
    ps = r.uniform(n_pla, 1.0, 10.0)
    prot = rotation_to(r.unit_vectors(n_pla))
    return {
        "mesh": (V, F),
        "spheres": (np.ascontiguousarray(sc), sr),
        "cylinders": (np.ascontiguousarray(cc), cr, cr.copy(), ch, crot),
        "planes": (np.ascontiguousarray(pc), ps, ps.copy(), prot),
    }


def mixed_rays(n: int = 8 * 1024 * 1024, seed: int = 13):
    return incoherent_rays(n, seed, 22.0)


def torus(n_tri_target: int, seed: int = 0, R: float = 1.0, r: float = 0.35, displacement: float = 0.05):
    """Closed torus mesh with about n_tri_target triangles (an nu x nv grid, 2 triangles per cell), each vertex
    moved along its tube normal by displacement * r * N(0,1)."""
    cells = max(8, n_tri_target // 2)
    nv = max(3, int(round(math.sqrt(cells * r / R))))
    nu = max(3, int(round(cells / nv)))
    u = 2 * math.pi * np.arange(nu) / nu
    v = 2 * math.pi * np.arange(nv) / nv
    U, Vv = np.meshgrid(u, v, indexing="ij")
    g = Rng(seed).normal(nu * nv).reshape(nu, nv)
    rr = r * (1.0 + displacement * g)
    x = (R + rr * np.cos(Vv)) * np.cos(U)
    y = (R + rr * np.cos(Vv)) * np.sin(U)
    z = rr * np.sin(Vv)
    verts = np.stack([x, y, z], axis=-1).reshape(-1, 3)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    faces = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return np.ascontiguousarray(verts), np.ascontiguousarray(faces.astype(np.uint32))


def uneven_batch(count: int = 32, seed: int = 41, tri_range=(1_000, 500_000), ray_range=(16_384, 2_097_152)):
    """Batched uneven workload: `count` triangle scenes (displaced icospheres and tori alternating, sizes
    log-uniform in tri_range) each with its own incoherent rays (count log-uniform in ray_range, origins in the
    scene's bounding cube scaled by 1.5). Returns a list of dicts with vertices, faces, origins, directions."""
    rng = Rng(seed)
    lt = np.exp(rng.uniform(count, math.log(tri_range[0]), math.log(tri_range[1])))
    lr = np.exp(rng.uniform(count, math.log(ray_range[0]), math.log(ray_range[1])))
    out = []
    for k in range(count):
        target = int(lt[k])
        if k % 2 == 0:
            level = max(1, int(round(math.log(max(target, 20) / 20.0, 4))))
            V, F = icosphere(level, seed=seed + k)
        else:
            V, F = torus(target, seed=seed + k)
        o, d = incoherent_rays(int(lr[k]), seed=seed + 1000 + k, box=1.5)
        out.append({"vertices": V, "faces": F, "origins": o, "directions": d})
    return out


def box_mesh(half: float):
    """Closed axis-aligned cube [-half, half]^3 as 8 vertices and 12 triangles."""
    v = np.array([[x, y, z] for x in (-half, half) for y in (-half, half) for z in (-half, half)], dtype=np.float64)
    f = np.array([[0, 1, 3], [0, 3, 2], [4, 6, 7], [4, 7, 5],   # x = -h, x = +h
                  [0, 4, 5], [0, 5, 1], [2, 3, 7], [2, 7, 6],   # y = -h, y = +h
                  [0, 2, 6], [0, 6, 4], [1, 5, 7], [1, 7, 3]],  # z = -h, z = +h
                 dtype=np.uint32)
    return v, f


def mirror_box_scene(level: int = 8, half: float = 3.0, seed: int = 7):
    """Displaced icosphere inside a 12-triangle box, merged into one mesh (box vertices and faces last)."""
    sv, sf = icosphere(level, seed=seed)
    bv, bf = box_mesh(half)
    v = np.concatenate([sv, bv])
    f = np.concatenate([sf, bf + len(sv)]).astype(np.uint32)
    return np.ascontiguousarray(v), np.ascontiguousarray(f)


def mirror_rays(n: int = 4 * 1024 * 1024, seed: int = 21):
    """Incoherent rays for mirror tracing: origins uniform in [-2, 2]^3, directions uniform on the sphere."""
    return incoherent_rays(n, seed, 2.0)


def mirror_weights(n: int, bounces: int = 3, seed: int = 29):
    """Per-ray, per-bounce loss weights w ~ U(0, 1), shape (bounces, n)."""
    return Rng(seed).uniform(bounces * n).reshape(bounces, n)


def scene_primitives():
    """examples/scene_primitives/main.cc:47-81 (two cylinders, then two spheres)."""
    rot = np.array([0, 0, 1, 0, 1, 0, -1, 0, 0], float)
    cyl = (np.array([[-3.0, 3, 0], [1, 4, 0]]), np.array([1.0, 0.5]), np.array([1.0, 0.5]), np.array([1.0, 2.0]),
           np.stack([rot, rot]))
    sph = (np.array([[1.0, 1, 0], [-2, 10, 0]]), np.array([0.5, 1.5]))
    return {"cylinders": cyl, "spheres": sph}


def scene_primitives_rays(width: int = 1080, height: int = 720):
    """examples/scene_primitives/main.cc:89-92: origin (0,5,20), raw direction (x/W - 0.5, y/H - 0.5, -1)."""
    ys, xs = np.meshgrid(np.arange(height), np.arange(width), indexing="ij")
    d = np.stack([xs.reshape(-1) / float(width) - 0.5, ys.reshape(-1) / float(height) - 0.5,
                  -np.ones(width * height)], axis=1)
    o = np.broadcast_to(np.array([0.0, 5.0, 20.0]), d.shape).copy()
    return o, np.ascontiguousarray(d)
