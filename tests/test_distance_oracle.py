"""Pins the CPU oracle of distance_to_surface (oracle/distance_oracle.c).

* Cylinders: all 207 active known answers of the reference's own test (test/test_primitives/test_cylinder.cpp,
  SUBCASE "distance to surface"; fixture tests/golden/cylinder_distance_cases.json) under its doctest Approx.
* Spheres: points placed at a known signed offset along a direction from the center.
* Planes: points placed at a known height above a rotated rectangle centered at x = y = 0 (where the reference's
  comparison of local coordinates with the world-space center is the geometric test), points beside it, and the
  reference's behaviour for an off-origin center pinned on explicit cases.
* Triangles: the closest point against an independent construction (the best of the plane projection when it
  falls inside, and the three clamped segment projections), on random triangles and points in every region.
* Scene query: the brute-force closest surface against per-pair distances, with ties across types and ids.
"""
import json
import os

import numpy as np
import pytest

from blazert_amd import workloads as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DBL_MAX = np.finfo(np.float64).max


def test_cylinder_known_answers(oracle_mod):
    doc = json.load(open(os.path.join(GOLDEN, "cylinder_distance_cases.json")))
    eps = doc["approx_epsilon"]
    assert len(doc["cases"]) == 207
    for c in doc["cases"]:
        got = oracle_mod.cylinder_distance(c["center"], c["a"], c["b"], c["height"], c["rotation"], c["point"])
        exp = c["distance"]
        assert abs(got - exp) <= eps * (1 + max(abs(got), abs(exp))), (c["line"], c["point"], exp, got)


def test_sphere_offsets(oracle_mod):
    r = W.Rng(3)
    for i in range(200):
        c = r.uniform(3, -5, 5)
        rad = float(r.uniform(1, 0.1, 3)[0])
        u = r.unit_vectors(1)[0]
        off = float(r.uniform(1, -0.9, 2.0)[0]) * rad
        p = c + (rad + off) * u
        assert abs(oracle_mod.sphere_distance(c, rad, p) - abs(off)) <= 1e-13 * (1 + np.abs(c).max() + rad)
    assert oracle_mod.sphere_distance([1.0, 2.0, 3.0], 0.5, [1.0, 2.0, 3.0]) == 0.5


def test_plane_heights_and_reference_offset_behaviour(oracle_mod):
    r = W.Rng(4)
    for i in range(200):
        R = W.rotation_to(r.unit_vectors(1))[0].reshape(3, 3)
        c = np.array([0.0, 0.0, float(r.uniform(1, -3, 3)[0])])
        dx, dy = (float(x) for x in r.uniform(2, 0.5, 4.0))
        lx, ly = float(r.uniform(1, -0.49, 0.49)[0]) * dx, float(r.uniform(1, -0.49, 0.49)[0]) * dy
        h = float(r.uniform(1, -2, 2)[0])
        p = c + R @ np.array([lx, ly, h])
        got = oracle_mod.plane_distance(c, dx, dy, R.reshape(-1), p)
        assert abs(got - abs(h)) <= 1e-13 * 10, (got, h)
        beside = c + R @ np.array([0.6 * dx + abs(lx), ly, h])
        assert oracle_mod.plane_distance(c, dx, dy, R.reshape(-1), beside) == DBL_MAX
    # plane.h:272-284 tests the local coordinates against the world-space center +- half extent: for a plane
    # centered at x = 10 a point straight above the center is rejected and a point 10 units beside it is accepted
    I = np.eye(3).reshape(-1)
    assert oracle_mod.plane_distance([10.0, 0.0, 0.0], 2.0, 2.0, I, [10.0, 0.0, 3.0]) == DBL_MAX
    assert oracle_mod.plane_distance([10.0, 0.0, 0.0], 2.0, 2.0, I, [20.0, 0.0, 3.0]) == 3.0


def _closest_independent(v0, v1, v2, p):
    """Closest point on a triangle: the plane projection if its barycentrics are inside, else the best clamped
    projection onto the three edges (a different construction from Eberly's region method)."""
    n = np.cross(v1 - v0, v2 - v0)
    q = p - np.dot(p - v0, n) / np.dot(n, n) * n
    # barycentrics of q by areas
    a0 = np.dot(np.cross(v1 - q, v2 - q), n)
    a1 = np.dot(np.cross(v2 - q, v0 - q), n)
    a2 = np.dot(np.cross(v0 - q, v1 - q), n)
    if a0 >= 0 and a1 >= 0 and a2 >= 0:
        return q
    best, bd = None, np.inf
    for a, b in ((v0, v1), (v1, v2), (v2, v0)):
        t = np.clip(np.dot(p - a, b - a) / np.dot(b - a, b - a), 0.0, 1.0)
        x = a + t * (b - a)
        d = np.dot(p - x, p - x)
        if d < bd:
            best, bd = x, d
    return best


def test_triangle_closest_point_independent(oracle_mod):
    r = W.Rng(5)
    worst = 0.0
    for i in range(3000):
        v = r.uniform(9, -2, 2).reshape(3, 3)
        n = np.cross(v[1] - v[0], v[2] - v[0])
        if np.linalg.norm(n) < 0.3:  # the reference's method is for well-formed triangles
            continue
        # points around the triangle: near each vertex, edge and the face, on both sides of the plane
        w = r.uniform(3, -0.6, 1.6)
        p = w[0] * v[0] + w[1] * v[1] + (1 - w[0] - w[1]) * v[2] + float(r.uniform(1, -1.5, 1.5)[0]) * n
        q = oracle_mod.closest_point_on_triangle(v[0], v[1], v[2], p)
        qi = _closest_independent(v[0], v[1], v[2], p)
        worst = max(worst, float(np.abs(q - qi).max()))
        d = oracle_mod.triangle_distance(v[0], v[1], v[2], p)
        assert abs(d - np.linalg.norm(p - qi)) <= 1e-12, (i, d, np.linalg.norm(p - qi))
    assert worst <= 1e-12, worst
    # closed forms: above the face, beyond a vertex, beside an edge
    v0, v1, v2 = np.array([0.0, 0, 0]), np.array([1.0, 0, 0]), np.array([0.0, 1, 0])
    assert oracle_mod.triangle_distance(v0, v1, v2, [0.25, 0.25, 2.0]) == 2.0
    assert oracle_mod.triangle_distance(v0, v1, v2, [-3.0, -4.0, 0.0]) == 5.0
    assert oracle_mod.triangle_distance(v0, v1, v2, [0.5, -2.0, 0.0]) == 2.0


def _mixed_small():
    m = W.mixed_scene(3000, 300, 60, 20, seed=8)
    return m


def _fill(m):
    def fill(s):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    return fill


def test_closest_surface_is_the_minimum_over_pairs(oracle_mod):
    m = _mixed_small()
    s = oracle_mod.Scene()
    _fill(m)(s)
    s.commit()
    pts = W.incoherent_rays(300, 9, 22.0)[0]
    got = s.closest_surface(pts)
    counts = {0: len(m["mesh"][1]), 1: len(m["spheres"][1]), 3: len(m["cylinders"][1]), 2: len(m["planes"][1])}
    gid = {0: 0, 1: 1, 3: 2, 2: 3}  # geometry ids follow add order: mesh, spheres, cylinders, planes
    best = np.full(len(pts), np.inf)
    bp = np.full(len(pts), 0xFFFFFFFF, np.uint32)
    bg = np.full(len(pts), 0xFFFFFFFF, np.uint32)
    for t in (0, 1, 2, 3):  # type order decides ties
        n = counts[t]
        for k in range(n):
            d = s.distance_to_surface(t, pts, np.full(len(pts), k, np.uint32))
            better = d < best
            best = np.where(better, d, best)
            bp = np.where(better, k, bp)
            bg = np.where(better, gid[t], bg)
    assert np.array_equal(got["distance"].view(np.uint64), best.view(np.uint64))
    assert np.array_equal(got["prim_id"], bp) and np.array_equal(got["geom_id"], bg)


def test_closest_surface_ties(oracle_mod):
    """Equal distances: the earlier geometry type wins, then the lower primitive id."""
    s = oracle_mod.Scene()
    s.add_spheres(np.array([[0.0, 0, 0], [0.0, 0, 0], [5.0, 0, 0]]), [1.0, 1.0, 1.0])
    s.add_mesh(np.array([[-3.0, -1, 2], [3.0, -1, 2], [0.0, 3, 2]]), np.array([[0, 1, 2]], np.uint32))
    s.commit()
    # (0,0,1.5): sphere surface at 0.5 (prims 0 and 1 tie), triangle plane z = 2 at 0.5 -> triangle (type 0) wins
    out = s.closest_surface(np.array([[0.0, 0.0, 1.5], [0.0, 0.0, -1.5]]))
    assert out["distance"][0] == 0.5 and out["geom_id"][0] == 1 and out["prim_id"][0] == 0
    assert out["distance"][1] == 0.5 and out["geom_id"][1] == 0 and out["prim_id"][1] == 0
    # an id out of range gives NaN in the per-pair query
    d = s.distance_to_surface(1, np.zeros((1, 3)), np.array([7], np.uint32))
    assert np.isnan(d[0])


@pytest.mark.parametrize("type_", [0, 1, 2, 3])
def test_pair_batch_matches_single_calls(oracle_mod, type_):
    m = _mixed_small()
    s = oracle_mod.Scene()
    _fill(m)(s)
    s.commit()
    key = {0: "mesh", 1: "spheres", 2: "planes", 3: "cylinders"}[type_]
    n = len(m[key][1])
    pts = W.incoherent_rays(500, 11, 22.0)[0]
    ids = (W.Rng(12).uniform(500) * n).astype(np.uint32)
    got = s.distance_to_surface(type_, pts, ids)
    for i in range(0, 500, 37):
        k = int(ids[i])
        if type_ == 0:
            V, F = m["mesh"]
            ref = oracle_mod.triangle_distance(V[F[k, 0]], V[F[k, 1]], V[F[k, 2]], pts[i])
        elif type_ == 1:
            ref = oracle_mod.sphere_distance(m["spheres"][0][k], m["spheres"][1][k], pts[i])
        elif type_ == 2:
            c, dx, dy, R = m["planes"]
            ref = oracle_mod.plane_distance(c[k], dx[k], dy[k], R[k], pts[i])
        else:
            c, a, b, h, R = m["cylinders"]
            ref = oracle_mod.cylinder_distance(c[k], a[k], b[k], h[k], R[k], pts[i])
        assert got[i] == ref or (np.isnan(got[i]) and np.isnan(ref))
