"""Pins the CPU oracle against every known answer of the reference's primitive test suite.

tests/golden/reference_cases.json holds every assert_* of blazert's test/test_primitives/test_trimesh.cpp,
test_sphere.cpp, test_plane.cpp and test_cylinder.cpp (610 asserts; the script that extracts them,
tests/golden/make_reference_cases.py, walks the doctest SUBCASE tree). Each case is checked the way the
reference's assert helper checks it (test/test_primitives/assert_helper.h):

* intersect_primitive (:123-146): primitive 0 tested alone, then post_traversal; hit flag, prim_id, hit distance and
  normal under doctest::Approx (a miss leaves the RayHit defaults: prim_id -1, distance max, normal 0);
* traverse (:212-228): BVH built with the default options and traversed; the same fields;
* traverse_distance (:178-190) and traverse_collection (:156-173): hit flag and distance (and normal);
* bounding_box, bounding_box_collection, primitive_center (:28-95) and distance_to_surface (:109-113).

The reference instantiates every one of these tests for T = float as well (TEST_CASE_TEMPLATE(..., float,
double)), so the intersection cases also run on the FP32 oracle (the reference's Scene<float> path).
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "reference_cases.json")
DOC = json.load(open(GOLDEN))
EPS = DOC["approx_epsilon"]
TYPES = {"mesh": 0, "spheres": 1, "planes": 2, "cylinders": 3}
INVALID = 0xFFFFFFFF


def approx(a, b):
    """doctest::Approx with its defaults: |a - b| < eps * (scale + max(|a|, |b|)), scale 1."""
    a = float(a)
    b = float(b)
    return abs(a - b) < EPS * (1.0 + max(abs(a), abs(b)))


def rot9(m):
    return np.array(m, float).reshape(-1, 9)


def add_geometry(s, sc):
    g = sc["geometry"]
    if g == "mesh":
        return s.add_mesh(np.array(sc["vertices"], float).reshape(-1, 3), np.array(sc["faces"], np.uint32).reshape(-1, 3))
    if g == "spheres":
        return s.add_spheres(np.array(sc["centers"], float).reshape(-1, 3), np.array(sc["radii"], float))
    if g == "planes":
        return s.add_planes(np.array(sc["centers"], float).reshape(-1, 3), sc["dxs"], sc["dys"], rot9(sc["rotations"]))
    return s.add_cylinders(np.array(sc["centers"], float).reshape(-1, 3), sc["a"], sc["b"], sc["heights"],
                           rot9(sc["rotations"]))


def cases(*asserts):
    return [pytest.param(c, id=f"{c['ref'].split('/')[-1]}") for c in DOC["cases"] if c["assert"] in asserts]


def test_every_reference_assert_is_present():
    """The fixture covers every non-WARN assert of the four files (counts from the extractor's walk)."""
    n = {}
    for c in DOC["cases"]:
        k = (c["scene"]["geometry"], c["assert"])
        n[k] = n.get(k, 0) + 1
    assert n[("cylinders", "intersect_primitive")] == 88 and n[("cylinders", "traverse")] == 88
    assert n[("planes", "intersect_primitive")] == 26 and n[("planes", "traverse")] == 26
    assert n[("mesh", "intersect_primitive")] == 24 and n[("mesh", "traverse")] == 24
    assert n[("mesh", "traverse_distance")] == 32
    assert n[("spheres", "intersect_primitive")] == 6 and n[("spheres", "traverse")] == 6
    assert n[("cylinders", "distance_to_surface")] == 207 and n[("planes", "distance_to_surface")] == 40
    assert len(DOC["cases"]) == 610


def check_hit(case, hit, h, tmax, normal=True, prim=True):
    assert bool(hit) == case["hit"], (hit, case["hit"])
    if prim:
        assert h.prim_id == (case["prim_id"] & INVALID), (h.prim_id, case["prim_id"])
    if case["hit"]:
        assert approx(h.t, case["t"]), (h.t, case["t"])
    else:
        assert h.t == tmax
    if normal and "normal" in case:
        for k in range(3):
            assert approx(h.normal[k], case["normal"][k]), (list(h.normal), case["normal"])


# --------------------------------------------------------------------------------------------------- FP64
@pytest.mark.parametrize("case", cases("intersect_primitive", "traverse", "traverse_distance", "traverse_collection"))
def test_reference_intersection_fp64(oracle_mod, case):
    s = oracle_mod.Scene()
    add_geometry(s, case["scene"])
    s.commit()
    t = TYPES[case["scene"]["geometry"]]
    ray = oracle_mod.make_ray(case["org"], case["dir"])
    if case["assert"] == "intersect_primitive":
        hit, h = s.intersect_single(t, 0, ray)
    else:
        hit, h, _ = s.traverse(t, ray)
    check_hit(case, hit, h, np.finfo(np.float64).max, normal=case["assert"] != "traverse_distance",
              prim=case["assert"] in ("intersect_primitive", "traverse"))


@pytest.mark.parametrize("case", cases("bounding_box", "bounding_box_collection", "primitive_center"))
def test_reference_boxes_and_centers(oracle_mod, case):
    s = oracle_mod.Scene()
    add_geometry(s, case["scene"])
    s.commit()
    t = TYPES[case["scene"]["geometry"]]
    if case["assert"] == "primitive_center":
        got = s.prim_center(t, case["prim"])
        assert all(approx(got[k], case["center"][k]) for k in range(3)), (got, case["center"])
        return
    if case["assert"] == "bounding_box":
        mn, mx = s.prim_box(t, case["prim"])
    else:  # assert_helper.h:52-82: union of the primitive boxes, starting from (max, min)
        n = len(case["scene"]["faces"])
        boxes = [s.prim_box(t, i) for i in range(n)]
        mn = [min(b[0][k] for b in boxes) for k in range(3)]
        mx = [max(b[1][k] for b in boxes) for k in range(3)]
    for k in range(3):
        assert approx(mn[k], case["bmin"][k]) and approx(mx[k], case["bmax"][k]), (mn, mx, case)


@pytest.mark.parametrize("case", cases("distance_to_surface"))
def test_reference_distance_to_surface(oracle_mod, case):
    sc = case["scene"]
    i = case["prim"]
    p = case["point"]
    if sc["geometry"] == "spheres":
        got = oracle_mod.sphere_distance(sc["centers"][i], sc["radii"][i], p)
    elif sc["geometry"] == "planes":
        got = oracle_mod.plane_distance(sc["centers"][i], sc["dxs"][i], sc["dys"][i], rot9(sc["rotations"])[i], p)
    else:
        got = oracle_mod.cylinder_distance(sc["centers"][i], sc["a"][i], sc["b"][i], sc["heights"][i],
                                           rot9(sc["rotations"])[i], p)
    assert approx(got, case["distance"]), (got, case["distance"])


# --------------------------------------------------------------------------------------------------- FP32
@pytest.mark.parametrize("case", cases("intersect_primitive", "traverse", "traverse_distance", "traverse_collection"))
def test_reference_intersection_fp32(oracle_mod, case):
    s = oracle_mod.SceneF32()
    add_geometry(s, case["scene"])
    s.commit()
    t = TYPES[case["scene"]["geometry"]]
    ray = oracle_mod.make_ray_f32(case["org"], case["dir"])
    if case["assert"] == "intersect_primitive":
        hit, h = s.intersect_single(t, 0, ray)
    else:  # a scene with one geometry type: intersect1 is that type's traverse
        hit, h, _ = s.intersect1(ray)
    check_hit(case, hit, h, np.finfo(np.float32).max, normal=case["assert"] != "traverse_distance",
              prim=case["assert"] in ("intersect_primitive", "traverse"))
