"""ctypes binding of the CPU oracle (oracle/blazert_oracle.c). TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's ``cpu_baseline`` leg import this module, and only
as the checker or the timed CPU baseline; the product package ``blazert_amd`` never imports it.
The oracle restates blazeRT's ``intersect1`` path (blazert/scene.h:161-211 and the functions it calls) and its
``distance_to_surface`` functions (distance_oracle.c).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("BLAZERT_ORACLE_LIB") or os.path.join(HERE, "liboracle.so")

TRIANGLES, SPHERES, PLANES, CYLINDERS = 0, 1, 2, 3
INVALID = 0xFFFFFFFF
REF_STACK_CAPACITY = 28


class Ray(C.Structure):
    _fields_ = [("org", C.c_double * 3), ("dir", C.c_double * 3), ("inv", C.c_double * 3),
                ("sign", C.c_uint32 * 3), ("tmin", C.c_double), ("tmax", C.c_double), ("any_hit", C.c_int)]


class Hit(C.Structure):
    _fields_ = [("normal", C.c_double * 3), ("u", C.c_double), ("v", C.c_double), ("t", C.c_double),
                ("prim_id", C.c_uint32), ("geom_id", C.c_uint32)]

    @staticmethod
    def default() -> "Hit":
        h = Hit()
        h.t = np.finfo(np.float64).max
        h.prim_id = INVALID
        h.geom_id = INVALID
        return h


class Trace(C.Structure):
    _fields_ = [("node_visits", C.c_uint32), ("prim_tests", C.c_uint32), ("max_stack", C.c_uint32),
                ("overflow", C.c_uint32), ("tie", C.c_uint32), ("multi_cyl", C.c_uint32)]


TRACE_DTYPE = np.dtype([("node_visits", np.uint32), ("prim_tests", np.uint32), ("max_stack", np.uint32),
                        ("overflow", np.uint32), ("tie", np.uint32), ("multi_cyl", np.uint32)])


class Node(C.Structure):
    _fields_ = [("min", C.c_double * 3), ("max", C.c_double * 3), ("leaf", C.c_uint32), ("axis", C.c_uint32),
                ("children", C.c_uint32 * 2), ("prim_begin", C.c_uint32), ("prim_count", C.c_uint32)]


NODE_DTYPE = np.dtype([("min", np.float64, 3), ("max", np.float64, 3), ("leaf", np.uint32), ("axis", np.uint32),
                       ("children", np.uint32, 2), ("prim_begin", np.uint32), ("prim_count", np.uint32)])
assert NODE_DTYPE.itemsize == C.sizeof(Node)


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint32) for n in ("max_tree_depth", "leaf_nodes", "max_num_prims_per_leaf",
                                          "min_num_prims_per_leaf", "branch_nodes", "bad_splits")]


_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc). Returns the library path."""
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.c_void_p
        dp = C.POINTER(C.c_double)
        up = C.POINTER(C.c_uint32)
        L.orc_scene_new.restype = P
        L.orc_scene_new.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_scene_free.argtypes = [P]
        L.orc_add_mesh.restype = C.c_uint32
        L.orc_add_mesh.argtypes = [P, dp, C.c_uint32, up, C.c_uint32]
        L.orc_add_spheres.restype = C.c_uint32
        L.orc_add_spheres.argtypes = [P, dp, dp, C.c_uint32]
        L.orc_add_planes.restype = C.c_uint32
        L.orc_add_planes.argtypes = [P, dp, dp, dp, dp, C.c_uint32]
        L.orc_add_cylinders.restype = C.c_uint32
        L.orc_add_cylinders.argtypes = [P, dp, dp, dp, dp, dp, C.c_uint32]
        L.orc_commit.argtypes = [P]
        L.orc_prim_box.argtypes = [P, C.c_int, C.c_uint32, dp, dp]
        L.orc_prim_center.argtypes = [P, C.c_int, C.c_uint32, dp]
        L.orc_tree_size.restype = C.c_uint32
        L.orc_tree_size.argtypes = [P, C.c_int]
        L.orc_tree_export.argtypes = [P, C.c_int, C.c_void_p, up, C.POINTER(Stats)]
        L.orc_collection_scene.restype = P
        L.orc_collection_scene.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, dp, dp, dp, C.c_uint32]
        L.orc_ray_init.argtypes = [C.POINTER(Ray), dp, dp, C.c_double, C.c_double, C.c_int]
        L.orc_intersect_node.argtypes = [dp, dp, dp, dp, C.POINTER(Ray)]
        L.orc_intersect_single.argtypes = [P, C.c_int, C.c_uint32, C.POINTER(Ray), C.POINTER(Hit)]
        L.orc_traverse.argtypes = [P, C.c_int, C.POINTER(Ray), C.POINTER(Hit), C.POINTER(Trace)]
        L.orc_intersect1.argtypes = [P, C.POINTER(Ray), C.POINTER(Hit), C.POINTER(Trace)]
        L.orc_intersect_batch.restype = C.c_uint64
        L.orc_intersect_batch.argtypes = [P, C.c_uint64] + [dp] * 8 + [dp, dp, dp, up, up, dp, dp, dp, C.c_void_p,
                                                                        C.c_int, C.c_int]
        L.orc_bruteforce_type.argtypes = [P, C.c_int, C.POINTER(Ray), C.POINTER(Hit)]
        L.orc_bruteforce_batch.restype = C.c_uint64
        L.orc_bruteforce_batch.argtypes = [P, C.c_int, C.c_uint64] + [dp] * 8 + [dp, up, C.c_int]
        L.orc_mirror_trace.restype = None
        L.orc_mirror_trace.argtypes = [P, dp, up, C.c_uint64] + [dp] * 8 + [C.c_uint32, C.c_double, dp, dp, dp, up] + \
            [dp] * 6 + [C.c_int]
        L.orc_mirror_loss.restype = None
        L.orc_mirror_loss.argtypes = [dp, up, C.c_uint64] + [dp] * 6 + [C.c_uint32, up] + [dp] * 5 + [C.c_int]
        L.orc_prim_params.argtypes = [P, C.c_int, C.c_uint32, dp]
        L.orc_prim_count.restype = C.c_uint32
        L.orc_geom_id.restype = C.c_uint32
        L.orc_geom_id.argtypes = [P, C.c_int]
        L.orc_prim_count.argtypes = [P, C.c_int]
        for name, nargs in (("orc_sphere_distance", [dp, C.c_double, dp]),
                            ("orc_plane_distance", [dp, C.c_double, C.c_double, dp, dp]),
                            ("orc_cylinder_distance", [dp, C.c_double, C.c_double, C.c_double, dp, dp]),
                            ("orc_triangle_distance", [dp, dp, dp, dp])):
            getattr(L, name).restype = C.c_double
            getattr(L, name).argtypes = nargs
        L.orc_closest_point_on_triangle.restype = None
        L.orc_closest_point_on_triangle.argtypes = [dp, dp, dp, dp, dp]
        L.orc_distance_batch.restype = None
        L.orc_distance_batch.argtypes = [P, C.c_int, C.c_uint64, dp, dp, dp, up, dp, C.c_int]
        L.orc_closest_batch.restype = None
        L.orc_closest_batch.argtypes = [P, C.c_uint64, dp, dp, dp, dp, up, up, C.c_int]
        L.orc_mirror_grad.restype = None
        L.orc_mirror_grad.argtypes = [dp, up, C.c_uint64] + [dp] * 6 + [C.c_uint32, up] + [dp] * 7 + [C.c_int]
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def _up(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32)) if a is not None else None


def _c3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def make_ray(org, direction, tmin=0.0, tmax=None, any_hit=False) -> Ray:
    r = Ray()
    if tmax is None:
        tmax = np.finfo(np.float64).max
    lib().orc_ray_init(C.byref(r), _c3(org), _c3(direction), float(tmin), float(tmax), int(any_hit))
    return r


def intersect_node(bmin, bmax, ray: Ray, tmin: float, tmax: float):
    a = C.c_double(tmin)
    b = C.c_double(tmax)
    hit = lib().orc_intersect_node(C.byref(a), C.byref(b), _c3(bmin), _c3(bmax), C.byref(ray))
    return bool(hit), a.value, b.value


class Scene:
    """Oracle scene with blazeRT's add/commit/intersect1 semantics."""

    def __init__(self, min_leaf_primitives=4, max_tree_depth=28, bin_size=64):
        self._keep = []
        self.h = lib().orc_scene_new(min_leaf_primitives, max_tree_depth, bin_size)

    def __del__(self):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            try:
                lib().orc_scene_free(h)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass

    def _k(self, *arrs):
        self._keep.extend(arrs)

    def add_mesh(self, vertices, faces) -> int:
        v = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        self._k(v, f)
        return lib().orc_add_mesh(self.h, _dp(v), len(v), _up(f), len(f))

    def add_spheres(self, centers, radii) -> int:
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        r = np.ascontiguousarray(radii, dtype=np.float64).reshape(-1)
        return lib().orc_add_spheres(self.h, _dp(c), _dp(r), len(r))

    def add_planes(self, centers, dxs, dys, rotations) -> int:
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        dx = np.ascontiguousarray(dxs, dtype=np.float64).reshape(-1)
        dy = np.ascontiguousarray(dys, dtype=np.float64).reshape(-1)
        R = np.ascontiguousarray(rotations, dtype=np.float64).reshape(-1, 9)
        return lib().orc_add_planes(self.h, _dp(c), _dp(dx), _dp(dy), _dp(R), len(dx))

    def add_cylinders(self, centers, semi_axes_a, semi_axes_b, heights, rotations) -> int:
        c = np.ascontiguousarray(centers, dtype=np.float64).reshape(-1, 3)
        a = np.ascontiguousarray(semi_axes_a, dtype=np.float64).reshape(-1)
        b = np.ascontiguousarray(semi_axes_b, dtype=np.float64).reshape(-1)
        h = np.ascontiguousarray(heights, dtype=np.float64).reshape(-1)
        R = np.ascontiguousarray(rotations, dtype=np.float64).reshape(-1, 9)
        return lib().orc_add_cylinders(self.h, _dp(c), _dp(a), _dp(b), _dp(h), _dp(R), len(a))

    def commit(self) -> bool:
        return bool(lib().orc_commit(self.h))

    def prim_box(self, type_, prim):
        mn = np.zeros(3)
        mx = np.zeros(3)
        rc = lib().orc_prim_box(self.h, type_, prim, _dp(mn), _dp(mx))
        if rc:
            raise IndexError(prim)
        return mn, mx

    def prim_center(self, type_, prim):
        c = np.zeros(3)
        if lib().orc_prim_center(self.h, type_, prim, _dp(c)):
            raise IndexError(prim)
        return c

    def tree(self, type_):
        n = lib().orc_tree_size(self.h, type_)
        nodes = np.zeros(n, dtype=NODE_DTYPE)
        nprim = 0 if n == 0 else None
        st = Stats()
        # prim_order length = collection size = sum of leaf counts
        lib().orc_tree_export(self.h, type_, nodes.ctypes.data_as(C.c_void_p), None, C.byref(st))
        nprim = int(nodes["prim_count"][nodes["leaf"] == 1].sum())
        order = np.zeros(max(nprim, 1), dtype=np.uint32)
        lib().orc_tree_export(self.h, type_, nodes.ctypes.data_as(C.c_void_p), _up(order), C.byref(st))
        stats = {name: getattr(st, name) for name, _ in Stats._fields_}
        return nodes, order[:nprim], stats

    def intersect_single(self, type_, prim, ray: Ray):
        h = Hit.default()
        hit = lib().orc_intersect_single(self.h, type_, prim, C.byref(ray), C.byref(h))
        return bool(hit), h

    def traverse(self, type_, ray: Ray):
        h = Hit.default()
        tr = Trace()
        hit = lib().orc_traverse(self.h, type_, C.byref(ray), C.byref(h), C.byref(tr))
        return bool(hit), h, tr

    def intersect1(self, ray: Ray, hit: Hit | None = None):
        h = hit if hit is not None else Hit.default()
        tr = Trace()
        res = lib().orc_intersect1(self.h, C.byref(ray), C.byref(h), C.byref(tr))
        return bool(res), h, tr

    def bruteforce(self, type_, ray: Ray):
        h = Hit.default()
        hit = lib().orc_bruteforce_type(self.h, type_, C.byref(ray), C.byref(h))
        return bool(hit), h

    def bruteforce_batch(self, type_, org, dirs, tmin=None, tmax=None, nthreads=0):
        """Closest accepted primitive of one type per ray, every primitive tested (no BVH): (t, user prim id)."""
        org = np.ascontiguousarray(org, dtype=np.float64).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = len(org)
        cols = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k])
                                                                       for k in range(3)]
        tmn = None if tmin is None else np.ascontiguousarray(np.broadcast_to(tmin, (n,)), dtype=np.float64)
        tmx = None if tmax is None else np.ascontiguousarray(np.broadcast_to(tmax, (n,)), dtype=np.float64)
        t = np.zeros(n)
        prim = np.zeros(n, dtype=np.uint32)
        lib().orc_bruteforce_batch(self.h, type_, n, *[_dp(c) for c in cols], _dp(tmn), _dp(tmx), _dp(t), _up(prim),
                                   int(nthreads))
        return t, prim

    def bruteforce_merged(self, org, dirs, types=(TRIANGLES, SPHERES, PLANES), tmin=None, tmax=None, nthreads=0):
        """blazert/scene.h:161-211 merge over per-type brute-force hits, in the reference's type order: the first type
        present is copied, later ones replace it only when strictly closer. Returns (t, prim_id, geom_id)."""
        n = len(np.asarray(org).reshape(-1, 3))
        t = np.full(n, np.finfo(np.float64).max)
        prim = np.full(n, INVALID, np.uint32)
        geom = np.full(n, INVALID, np.uint32)
        for ty in types:
            if lib().orc_prim_count(self.h, ty) == 0:
                continue
            tt, pp = self.bruteforce_batch(ty, org, dirs, tmin, tmax, nthreads)
            take = (pp != INVALID) & (tt < t)
            t[take], prim[take], geom[take] = tt[take], pp[take], lib().orc_geom_id(self.h, ty)
        return t, prim, geom

    def intersect_batch(self, org, dirs, tmin=None, tmax=None, normals=True, trace=False, nthreads=0, any_hit=False):
        """Batched intersect1 on the CPU. org/dirs: (n,3) float64. Returns a dict of numpy arrays."""
        org = np.ascontiguousarray(org, dtype=np.float64).reshape(-1, 3)
        dirs = np.ascontiguousarray(dirs, dtype=np.float64).reshape(-1, 3)
        n = len(org)
        cols = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k])
                                                                       for k in range(3)]
        tmn = None if tmin is None else np.ascontiguousarray(np.broadcast_to(tmin, (n,)), dtype=np.float64)
        tmx = None if tmax is None else np.ascontiguousarray(np.broadcast_to(tmax, (n,)), dtype=np.float64)
        out = {k: np.zeros(n) for k in ("t", "u", "v")}
        out["prim_id"] = np.zeros(n, dtype=np.uint32)
        out["geom_id"] = np.zeros(n, dtype=np.uint32)
        nrm = [np.zeros(n) for _ in range(3)] if normals else [None, None, None]
        tr = np.zeros(n, dtype=TRACE_DTYPE) if trace else None
        lib().orc_intersect_batch(self.h, n, *[_dp(c) for c in cols], _dp(tmn), _dp(tmx), _dp(out["t"]),
                                  _dp(out["u"]), _dp(out["v"]), _up(out["prim_id"]), _up(out["geom_id"]),
                                  *[_dp(c) for c in nrm], tr.ctypes.data_as(C.c_void_p) if tr is not None else None,
                                  int(nthreads), int(any_hit))
        if normals:
            out["normal"] = np.stack(nrm, axis=1)
        if trace:
            out["trace"] = tr
        return out


    # ------------------------------------------------------------------------------------ distance_to_surface
    def distance_to_surface(self, type_, points, prims, nthreads=0):
        """Distance from point i to primitive prims[i] (user id) of one geometry type (distance_oracle.c)."""
        px, py, pz = _cols(points, len(prims))
        ids = np.ascontiguousarray(prims, dtype=np.uint32)
        out = np.zeros(len(ids))
        lib().orc_distance_batch(self.h, type_, len(ids), _dp(px), _dp(py), _dp(pz), _up(ids), _dp(out),
                                 int(nthreads))
        return out

    def closest_surface(self, points, nthreads=0):
        """Brute-force closest surface over the whole scene: distance, primitive id, geometry id."""
        points = np.asarray(points, dtype=np.float64).reshape(-1, 3)
        n = len(points)
        px, py, pz = _cols(points, n)
        out = {"distance": np.zeros(n), "prim_id": np.zeros(n, np.uint32), "geom_id": np.zeros(n, np.uint32)}
        lib().orc_closest_batch(self.h, n, _dp(px), _dp(py), _dp(pz), _dp(out["distance"]), _up(out["prim_id"]),
                                _up(out["geom_id"]), int(nthreads))
        return out


def classify_split_disagreements(scene: "Scene", org, dirs, idx, split_t, split_geom, cyl_geom=None, tmin=None,
                                 tmax=None) -> dict:
    """Explains the rays where the split-precision path (FP32 traversal, FP64 hit) and the FP64 reference disagree.

    The split path tests every primitive the reference's FP64 traversal tests, plus possibly some in nodes the
    reference culls; the primitive tests, hit distances and merge are FP64. So a disagreement can only be:
    * a tie: a runner-up within 1e-12 relative of the reference's t in its traversal, or a split hit at the same
      distance to 1e-12 (a candidate the reference never tested, within the -4 eps barycentric slack);
    * a cylinder: either side's hit is a cylinder, or two cylinders accept the ray. The reference accepts a cylinder
      without a distance check (last visited wins, blazert/primitives/cylinder.h:181-495), and bounds a rotated
      cylinder by a box from 6 axis points that can miss part of it (cylinder.h:115-147);
    * a reference stack overflow (28 slots, undefined behaviour in the reference).
    idx are the disagreeing ray indices into org/dirs; split_t / split_geom the split path's t and geom_id on them;
    cyl_geom the geometry id of the scene's cylinders. Returns counts and the indices nothing explains."""
    idx = np.asarray(idx, dtype=np.int64)
    if len(idx) == 0:
        return {"disagree": 0, "tie": 0, "cylinder": 0, "overflow": 0, "unexplained": 0,
                "unexplained_idx": np.zeros(0, np.int64)}
    sel = lambda a: None if a is None else np.broadcast_to(np.asarray(a, np.float64), (len(org),))[idx]
    ref = scene.intersect_batch(np.asarray(org)[idx], np.asarray(dirs)[idx], tmin=sel(tmin), tmax=sel(tmax),
                                normals=False, trace=True)
    tr = ref["trace"]
    ot = np.asarray(split_t, np.float64)
    og = np.asarray(split_geom).view(np.uint32) if np.asarray(split_geom).dtype.itemsize == 4 else \
        np.asarray(split_geom, np.uint32)
    tie = (tr["tie"] != 0) | (np.abs(ot - ref["t"]) <= 1e-12 * np.abs(ref["t"]))
    cyl = tr["multi_cyl"] != 0
    if cyl_geom is not None:
        cyl |= (ref["geom_id"] == cyl_geom) | (og == cyl_geom)
    ovf = tr["overflow"] != 0
    un = ~tie & ~cyl & ~ovf
    return {"disagree": int(len(idx)), "tie": int(tie.sum()), "cylinder": int((cyl & ~tie).sum()),
            "overflow": int((ovf & ~tie & ~cyl).sum()), "unexplained": int(un.sum()), "unexplained_idx": idx[un]}


def _v3(v):
    return np.ascontiguousarray(v, dtype=np.float64).reshape(-1)


def sphere_distance(c, r, p) -> float:
    return lib().orc_sphere_distance(_dp(_v3(c)), float(r), _dp(_v3(p)))


def plane_distance(c, dx, dy, R, p) -> float:
    return lib().orc_plane_distance(_dp(_v3(c)), float(dx), float(dy), _dp(_v3(R)), _dp(_v3(p)))


def cylinder_distance(c, a, b, h, R, p) -> float:
    return lib().orc_cylinder_distance(_dp(_v3(c)), float(a), float(b), float(h), _dp(_v3(R)), _dp(_v3(p)))


def triangle_distance(v0, v1, v2, p) -> float:
    return lib().orc_triangle_distance(_dp(_v3(v0)), _dp(_v3(v1)), _dp(_v3(v2)), _dp(_v3(p)))


def closest_point_on_triangle(v0, v1, v2, p):
    q = np.zeros(3)
    lib().orc_closest_point_on_triangle(_dp(_v3(v0)), _dp(_v3(v1)), _dp(_v3(v2)), _dp(_v3(p)), _dp(q))
    return q


# ---------------------------------------------------------------------------------------------- mirror tracing
def _cols(a, n):
    a = np.asarray(a, dtype=np.float64).reshape(n, 3)
    return [np.ascontiguousarray(a[:, k]) for k in range(3)]


def _opt(a, shape):
    return None if a is None else np.ascontiguousarray(np.broadcast_to(a, shape), dtype=np.float64)


def mirror_trace(scene: Scene, vertices, faces, org, dirs, bounces, bounce_tmin, tmin=None, tmax=None, nthreads=0):
    """Multi-bounce mirror tracing on the CPU (mirror_oracle.c). Returns t, u, v, prim_id (bounces, n) and
    position, direction (bounces, n, 3)."""
    V = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    n = len(np.asarray(org).reshape(-1, 3))
    cols = _cols(org, n) + _cols(dirs, n)
    tmn, tmx = _opt(tmin, (n,)), _opt(tmax, (n,))
    B = int(bounces)
    t, u, v = (np.zeros(B * n) for _ in range(3))
    prim = np.zeros(B * n, dtype=np.uint32)
    pr = [np.zeros(B * n) for _ in range(6)]
    lib().orc_mirror_trace(scene.h, _dp(V), _up(F), n, *[_dp(c) for c in cols], _dp(tmn), _dp(tmx), B,
                           float(bounce_tmin), _dp(t), _dp(u), _dp(v), _up(prim), *[_dp(c) for c in pr],
                           int(nthreads))
    return {"t": t.reshape(B, n), "u": u.reshape(B, n), "v": v.reshape(B, n), "prim_id": prim.reshape(B, n),
            "position": np.stack(pr[:3], axis=-1).reshape(B, n, 3),
            "direction": np.stack(pr[3:], axis=-1).reshape(B, n, 3)}


def _grad_inputs(org, dirs, prim_id, gt, gp):
    prim = np.ascontiguousarray(prim_id, dtype=np.uint32)
    B, n = prim.shape
    cols = _cols(org, n) + _cols(dirs, n)
    g_t = _opt(gt, (B, n))
    gps = [None, None, None] if gp is None else [np.ascontiguousarray(np.asarray(gp, dtype=np.float64)
                                                                      .reshape(B, n, 3)[..., k]) for k in range(3)]
    return prim, B, n, cols, g_t, gps


def mirror_loss(vertices, faces, org, dirs, prim_id, gt=None, gp=None, nthreads=0):
    """Per-ray loss sum_b gt[b]*t_b + gp[b].x_b of the bounce chain with the given (fixed) face ids."""
    V = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    prim, B, n, cols, g_t, gps = _grad_inputs(org, dirs, prim_id, gt, gp)
    loss = np.zeros(n)
    lib().orc_mirror_loss(_dp(V), _up(F), n, *[_dp(c) for c in cols], B, _up(prim), _dp(g_t),
                          *[_dp(c) for c in gps], _dp(loss), int(nthreads))
    return loss


def mirror_grad(vertices, faces, org, dirs, prim_id, gt=None, gp=None, nthreads=0):
    """Analytic FP64 gradient of sum(mirror_loss): (dL/dvertex (nv,3), dL/dorigin (n,3), dL/ddirection (n,3))."""
    V = np.ascontiguousarray(vertices, dtype=np.float64).reshape(-1, 3)
    F = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
    prim, B, n, cols, g_t, gps = _grad_inputs(org, dirs, prim_id, gt, gp)
    gv = np.zeros(V.size)
    go = np.zeros(3 * n)
    gd = np.zeros(3 * n)
    lib().orc_mirror_grad(_dp(V), _up(F), n, *[_dp(c) for c in cols], B, _up(prim), _dp(g_t),
                          *[_dp(c) for c in gps], _dp(gv), _dp(go), _dp(gd), int(nthreads))
    return gv.reshape(-1, 3), go.reshape(3, n).T.copy(), gd.reshape(3, n).T.copy()


# ---------------------------------------------------------------------------------------------- Scene<float>
# The same restatement compiled with -DORC_FLOAT (liboracle_f32.so): blazert::Scene<float>, every operation in FP32.
class RayF(C.Structure):
    _fields_ = [("org", C.c_float * 3), ("dir", C.c_float * 3), ("inv", C.c_float * 3),
                ("sign", C.c_uint32 * 3), ("tmin", C.c_float), ("tmax", C.c_float), ("any_hit", C.c_int)]


class HitF(C.Structure):
    _fields_ = [("normal", C.c_float * 3), ("u", C.c_float), ("v", C.c_float), ("t", C.c_float),
                ("prim_id", C.c_uint32), ("geom_id", C.c_uint32)]

    @staticmethod
    def default() -> "HitF":
        h = HitF()
        h.t = np.finfo(np.float32).max
        h.prim_id = INVALID
        h.geom_id = INVALID
        return h


NODE_DTYPE_F32 = np.dtype([("min", np.float32, 3), ("max", np.float32, 3), ("leaf", np.uint32), ("axis", np.uint32),
                           ("children", np.uint32, 2), ("prim_begin", np.uint32), ("prim_count", np.uint32)])
LIB_F32_PATH = os.environ.get("BLAZERT_ORACLE_F32_LIB") or os.path.join(HERE, "liboracle_f32.so")
_lib_f32 = None


def lib_f32():
    global _lib_f32
    if _lib_f32 is None:
        if not os.path.exists(LIB_F32_PATH):
            build(force=True)
        L = C.CDLL(LIB_F32_PATH)
        P = C.c_void_p
        fp = C.POINTER(C.c_float)
        up = C.POINTER(C.c_uint32)
        L.orc_scene_new.restype = P
        L.orc_scene_new.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32]
        L.orc_scene_free.argtypes = [P]
        L.orc_add_mesh.restype = C.c_uint32
        L.orc_add_mesh.argtypes = [P, fp, C.c_uint32, up, C.c_uint32]
        L.orc_add_spheres.restype = C.c_uint32
        L.orc_add_spheres.argtypes = [P, fp, fp, C.c_uint32]
        L.orc_add_planes.restype = C.c_uint32
        L.orc_add_planes.argtypes = [P, fp, fp, fp, fp, C.c_uint32]
        L.orc_add_cylinders.restype = C.c_uint32
        L.orc_add_cylinders.argtypes = [P, fp, fp, fp, fp, fp, C.c_uint32]
        L.orc_commit.argtypes = [P]
        L.orc_tree_size.restype = C.c_uint32
        L.orc_tree_size.argtypes = [P, C.c_int]
        L.orc_tree_export.argtypes = [P, C.c_int, C.c_void_p, up, C.POINTER(Stats)]
        L.orc_collection_scene.restype = P
        L.orc_collection_scene.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, fp, fp, fp, C.c_uint32]
        L.orc_ray_init.argtypes = [C.POINTER(RayF), fp, fp, C.c_float, C.c_float, C.c_int]
        L.orc_intersect_single.argtypes = [P, C.c_int, C.c_uint32, C.POINTER(RayF), C.POINTER(HitF)]
        L.orc_intersect1.argtypes = [P, C.POINTER(RayF), C.POINTER(HitF), C.POINTER(Trace)]
        L.orc_intersect_batch.restype = C.c_uint64
        L.orc_intersect_batch.argtypes = [P, C.c_uint64] + [fp] * 8 + [fp, fp, fp, up, up, fp, fp, fp, C.c_void_p,
                                                                        C.c_int, C.c_int]
        _lib_f32 = L
    return _lib_f32


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def make_ray_f32(org, direction, tmin=0.0, tmax=None, any_hit=False) -> RayF:
    r = RayF()
    if tmax is None:
        tmax = np.finfo(np.float32).max
    o = (C.c_float * 3)(*[float(np.float32(x)) for x in org])
    d = (C.c_float * 3)(*[float(np.float32(x)) for x in direction])
    lib_f32().orc_ray_init(C.byref(r), o, d, float(np.float32(tmin)), float(np.float32(tmax)), int(any_hit))
    return r


class SceneF32:
    """Oracle of blazert::Scene<float>: the add/commit/intersect1 path with T = float."""

    def __init__(self, min_leaf_primitives=4, max_tree_depth=28, bin_size=64):
        self._keep = []
        self.h = lib_f32().orc_scene_new(min_leaf_primitives, max_tree_depth, bin_size)

    def __del__(self):
        h = getattr(self, "h", None)
        self.h = None
        if h:
            try:
                lib_f32().orc_scene_free(h)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass

    @staticmethod
    def _f(a, cols=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a.reshape(-1, cols) if cols else a.reshape(-1)

    def add_mesh(self, vertices, faces) -> int:
        v = self._f(vertices, 3)
        f = np.ascontiguousarray(faces, dtype=np.uint32).reshape(-1, 3)
        self._keep.extend([v, f])
        return lib_f32().orc_add_mesh(self.h, _fp(v), len(v), _up(f), len(f))

    def add_spheres(self, centers, radii) -> int:
        c, r = self._f(centers, 3), self._f(radii)
        return lib_f32().orc_add_spheres(self.h, _fp(c), _fp(r), len(r))

    def add_planes(self, centers, dxs, dys, rotations) -> int:
        c, dx, dy, R = self._f(centers, 3), self._f(dxs), self._f(dys), self._f(rotations, 9)
        return lib_f32().orc_add_planes(self.h, _fp(c), _fp(dx), _fp(dy), _fp(R), len(dx))

    def add_cylinders(self, centers, semi_axes_a, semi_axes_b, heights, rotations) -> int:
        c, a, b = self._f(centers, 3), self._f(semi_axes_a), self._f(semi_axes_b)
        h, R = self._f(heights), self._f(rotations, 9)
        return lib_f32().orc_add_cylinders(self.h, _fp(c), _fp(a), _fp(b), _fp(h), _fp(R), len(a))

    def commit(self) -> bool:
        return bool(lib_f32().orc_commit(self.h))

    def tree(self, type_):
        n = lib_f32().orc_tree_size(self.h, type_)
        nodes = np.zeros(n, dtype=NODE_DTYPE_F32)
        st = Stats()
        lib_f32().orc_tree_export(self.h, type_, nodes.ctypes.data_as(C.c_void_p), None, C.byref(st))
        nprim = int(nodes["prim_count"][nodes["leaf"] == 1].sum())
        order = np.zeros(max(nprim, 1), dtype=np.uint32)
        lib_f32().orc_tree_export(self.h, type_, nodes.ctypes.data_as(C.c_void_p), _up(order), C.byref(st))
        return nodes, order[:nprim], {name: getattr(st, name) for name, _ in Stats._fields_}

    def intersect_single(self, type_, prim, ray: RayF):
        h = HitF.default()
        hit = lib_f32().orc_intersect_single(self.h, type_, prim, C.byref(ray), C.byref(h))
        return bool(hit), h

    def intersect1(self, ray: RayF, hit: HitF | None = None):
        h = hit if hit is not None else HitF.default()
        tr = Trace()
        res = lib_f32().orc_intersect1(self.h, C.byref(ray), C.byref(h), C.byref(tr))
        return bool(res), h, tr

    def intersect_batch(self, org, dirs, tmin=None, tmax=None, normals=True, trace=False, nthreads=0, any_hit=False):
        """Batched intersect1 with T = float: float32 in, float32 out."""
        org = self._f(org, 3)
        dirs = self._f(dirs, 3)
        n = len(org)
        cols = [np.ascontiguousarray(org[:, k]) for k in range(3)] + [np.ascontiguousarray(dirs[:, k])
                                                                       for k in range(3)]
        tmn = None if tmin is None else np.ascontiguousarray(np.broadcast_to(tmin, (n,)), dtype=np.float32)
        tmx = None if tmax is None else np.ascontiguousarray(np.broadcast_to(tmax, (n,)), dtype=np.float32)
        out = {k: np.zeros(n, np.float32) for k in ("t", "u", "v")}
        out["prim_id"] = np.zeros(n, dtype=np.uint32)
        out["geom_id"] = np.zeros(n, dtype=np.uint32)
        nrm = [np.zeros(n, np.float32) for _ in range(3)] if normals else [None, None, None]
        tr = np.zeros(n, dtype=TRACE_DTYPE) if trace else None
        lib_f32().orc_intersect_batch(self.h, n, *[_fp(c) for c in cols], _fp(tmn), _fp(tmx), _fp(out["t"]),
                                      _fp(out["u"]), _fp(out["v"]), _up(out["prim_id"]), _up(out["geom_id"]),
                                      *[_fp(c) for c in nrm],
                                      tr.ctypes.data_as(C.c_void_p) if tr is not None else None, int(nthreads),
                                      int(any_hit))
        if normals:
            out["normal"] = np.stack(nrm, axis=1)
        if trace:
            out["trace"] = tr
        return out


def collection_tree(box_min, box_max, centers, dtype=np.float64, min_leaf_primitives=4, max_tree_depth=28,
                    bin_size=64):
    """The reference builder (blazert/bvh/builder.h:48-82) on a caller-defined collection given by its boxes and
    centers ((n, 3) arrays), in T = dtype. Returns (nodes, leaf order, stats)."""
    f32 = np.dtype(dtype) == np.float32
    L = lib_f32() if f32 else lib()
    arrs = [np.ascontiguousarray(np.asarray(a, dtype).reshape(-1, 3)) for a in (box_min, box_max, centers)]
    ptr = _fp if f32 else _dp
    h = L.orc_collection_scene(min_leaf_primitives, max_tree_depth, bin_size, *[ptr(a) for a in arrs], len(arrs[0]))
    try:
        n = L.orc_tree_size(h, 1)
        nodes = np.zeros(n, dtype=NODE_DTYPE_F32 if f32 else NODE_DTYPE)
        order = np.zeros(max(len(arrs[0]), 1), dtype=np.uint32)
        st = Stats()
        L.orc_tree_export(h, 1, nodes.ctypes.data_as(C.c_void_p), _up(order), C.byref(st))
        return nodes, order[:len(arrs[0])], {name: getattr(st, name) for name, _ in Stats._fields_}
    finally:
        L.orc_scene_free(h)
