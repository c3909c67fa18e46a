"""Pins the CPU oracle against the reference's own known answers (tests/golden/known_answers.json).

Every case is transcribed from the blazert doctest suite (file:line in each case). The reference asserts
with doctest::Approx (relative 1.19e-5 with scale 1), so that is the tolerance used here.
"""
import json
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
DOC = json.load(open(GOLDEN))
EPS = DOC["approx_epsilon"]
TYPES = {"mesh": 0, "spheres": 1, "planes": 2, "cylinders": 3}


def approx(a, b):
    a = float(a)
    b = float(b)
    return abs(a - b) < EPS * (1.0 + max(abs(a), abs(b)))


def build_scene(oracle, case, order=None):
    s = oracle.Scene()
    ids = {}
    order = order or case.get("add_order") or [case["geometry"]]
    for g in order:
        if g == "mesh":
            m = case["mesh"]
            ids[g] = s.add_mesh(np.array(m["vertices"], float), np.array(m["faces"], np.uint32))
        elif g == "spheres":
            d = case["spheres"]
            ids[g] = s.add_spheres(np.array(d["centers"], float), np.array(d["radii"], float))
        elif g == "planes":
            d = case["planes"]
            ids[g] = s.add_planes(np.array(d["centers"], float), d["dxs"], d["dys"], np.array(d["rotations"], float))
        elif g == "cylinders":
            d = case["cylinders"]
            ids[g] = s.add_cylinders(np.array(d["centers"], float), d["a"], d["b"], d["heights"],
                                     np.array(d["rotations"], float))
    s.commit()
    return s, ids


def cases(kind):
    return [pytest.param(c, id=f"{c['name']}") for c in DOC["cases"] if c["kind"] == kind]


@pytest.mark.parametrize("case", cases("slab"))
def test_slab_known_answers(oracle_mod, case):
    ray = oracle_mod.make_ray(case["org"], case["dir"], case["tmin"], case["tmax"])
    hit, tmin, tmax = oracle_mod.intersect_node(case["bmin"], case["bmax"], ray, case["tmin"], case["tmax"])
    if case["severity"] == "warn":
        pytest.skip("WARN-only case in the reference")
    assert hit == case["hit"]
    assert approx(tmin, case["tmin_out"])
    assert approx(tmax, case["tmax_out"])


def _check_hit(case, hit, h):
    assert hit == case["hit"]
    if case["hit"]:
        if case.get("prim_id") is not None:
            assert h.prim_id == case["prim_id"]
        assert approx(h.t, case["t"])
        if case.get("normal") is not None:
            for k in range(3):
                assert approx(h.normal[k], case["normal"][k]), (list(h.normal), case["normal"])
    else:
        assert h.prim_id == 0xFFFFFFFF
        assert h.t == np.finfo(np.float64).max


@pytest.mark.parametrize("case", cases("primitive"))
@pytest.mark.parametrize("mode", ["single", "traverse"])
def test_primitive_known_answers(oracle_mod, case, mode):
    s, _ = build_scene(oracle_mod, case)
    ray = oracle_mod.make_ray(case["org"], case["dir"])
    t = TYPES[case["geometry"]]
    if mode == "single":
        hit, h = s.intersect_single(t, 0, ray)
    else:
        hit, h, _ = s.traverse(t, ray)
    _check_hit(case, hit, h)


@pytest.mark.parametrize("case", cases("traverse") + cases("traverse_distance"))
def test_traverse_known_answers(oracle_mod, case):
    s, _ = build_scene(oracle_mod, case)
    ray = oracle_mod.make_ray(case["org"], case["dir"])
    hit, h, _ = s.traverse(TYPES[case["geometry"]], ray)
    assert hit == case["hit"]
    assert approx(h.t, case["t"])
    if "normal" in case:
        for k in range(3):
            assert approx(h.normal[k], case["normal"][k])


@pytest.mark.parametrize("case", cases("box"))
def test_box_known_answers(oracle_mod, case):
    s, _ = build_scene(oracle_mod, case)
    mn, mx = s.prim_box(TYPES[case["geometry"]], case["prim"])
    for k in range(3):
        assert approx(mn[k], case["bmin"][k]) and approx(mx[k], case["bmax"][k]), (mn, mx)


@pytest.mark.parametrize("case", cases("center"))
def test_center_known_answers(oracle_mod, case):
    s, _ = build_scene(oracle_mod, case)
    c = s.prim_center(TYPES[case["geometry"]], case["prim"])
    for k in range(3):
        assert approx(c[k], case["center"][k])


@pytest.mark.parametrize("case", cases("scene"))
def test_scene_known_answers(oracle_mod, case):
    s, ids = build_scene(oracle_mod, case)
    for g, gid in case.get("expect_geom_ids", {}).items():
        assert ids[g] == gid
    ray = oracle_mod.make_ray(case["org"], case["dir"])
    hit, h, _ = s.intersect1(ray)
    assert hit == case["hit"]
    assert h.geom_id == case["geom_id"]
    assert h.prim_id == case["prim_id"]
    assert approx(h.t, case["t"])
    for k in range(3):
        assert approx(h.normal[k], case["normal"][k])


def test_add_twice_and_after_commit_return_invalid(oracle_mod):
    """blazert/scene.h:230-239: a second add of one type, or any add after commit, returns unsigned(-1)."""
    s = oracle_mod.Scene()
    assert s.add_spheres(np.zeros((1, 3)), [1.0]) == 0
    assert s.add_spheres(np.zeros((1, 3)), [1.0]) == oracle_mod.INVALID
    s.commit()
    v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    assert s.add_mesh(v, np.array([[0, 1, 2]], np.uint32)) == oracle_mod.INVALID
