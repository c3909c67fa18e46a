"""Python mirror of blazeRT's scene API over the C ABI.

``Scene`` follows ``blazert::BlazertScene<T>`` (blazert/scene.h:38-116): one add per geometry type, each add
returns the geometry id (or ``INVALID_ID`` = unsigned(-1) for a second add of a type or an add after commit),
``commit()`` builds the BVHs, ``intersect1(scene, ray, rayhit)`` is the single-ray query. ``Scene.intersect``
is the batched form over device-resident rays (torch tensors on the GPU; torch is used for device memory and
streams only, the traversal is the HIP code in libblazert_amd.so).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .capi import INVALID_ID, check, lib

DBL_MAX = float(np.finfo(np.float64).max)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


class Scene:
    """blazert::Scene<T> (blazert/blazert.h:17-20) on the MI355X: T = double by default, dtype=np.float32 for
    Scene<float> (FP32 geometry, build, traversal and hits; add / commit / intersect / intersect1 only)."""

    def __init__(self, min_leaf_primitives: int = 4, max_tree_depth: int = 28, bin_size: int = 64,
                 dtype=np.float64):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.dtype(np.float64), np.dtype(np.float32)):
            raise ValueError("Scene dtype must be float64 or float32")
        self.f32 = self.dtype == np.float32
        opts = capi.BuildOptions(min_leaf_primitives, max_tree_depth, bin_size)
        h = C.c_void_p()
        check((lib().brt_scene_create_f32 if self.f32 else lib().brt_scene_create)(C.byref(opts), C.byref(h)))
        self._h = h
        self.has_been_committed = False
        self.num_vertices = 0

    def close(self):
        h = getattr(self, "_h", None)
        self._h = None
        if h:
            try:
                lib().brt_scene_destroy(h)
            except Exception:  # interpreter shutdown: module globals may already be gone
                pass

    def __del__(self):
        self.close()

    @property
    def handle(self):
        return self._h

    def _add(self, fn, *args) -> int:
        gid = C.c_uint32(INVALID_ID)
        rc = fn(self._h, *args, C.byref(gid))
        if rc == capi.ERR_STATE:
            return INVALID_ID  # reference semantics: unsigned(-1), no exception (blazert/scene.h:237-239)
        check(rc)
        return gid.value

    def _a(self, a, cols=None):
        """Contiguous array of the scene type and its ctypes pointer."""
        x = np.ascontiguousarray(a, dtype=self.dtype)
        x = x.reshape(-1, cols) if cols else x.reshape(-1)
        return x

    def _p(self, x):
        return x.ctypes.data_as(C.POINTER(C.c_float if self.f32 else C.c_double))

    def _fn(self, name):
        return getattr(lib(), name + ("_f32" if self.f32 else ""))

    def add_mesh(self, vertices, triangles) -> int:
        v = self._a(vertices, 3)
        f = np.ascontiguousarray(triangles, dtype=np.uint32).reshape(-1, 3)
        gid = self._add(self._fn("brt_scene_add_mesh"), self._p(v), len(v), _up(f), len(f))
        if gid != INVALID_ID:
            self.num_vertices = len(v)
        return gid

    def add_spheres(self, centers, radii) -> int:
        c = self._a(centers, 3)
        r = self._a(radii)
        if len(c) != len(r):
            raise ValueError("centers and radii must have the same length")
        return self._add(self._fn("brt_scene_add_spheres"), self._p(c), self._p(r), len(r))

    def add_planes(self, centers, dxs, dys, rotations) -> int:
        c = self._a(centers, 3)
        dx, dy = self._a(dxs), self._a(dys)
        R = self._a(rotations, 9)
        if not (len(c) == len(dx) == len(dy) == len(R)):
            raise ValueError("plane arrays must have the same length")
        return self._add(self._fn("brt_scene_add_planes"), self._p(c), self._p(dx), self._p(dy), self._p(R), len(dx))

    def add_cylinders(self, centers, semi_axes_a, semi_axes_b, heights, rotations) -> int:
        c = self._a(centers, 3)
        a, b, h = self._a(semi_axes_a), self._a(semi_axes_b), self._a(heights)
        R = self._a(rotations, 9)
        if not (len(c) == len(a) == len(b) == len(h) == len(R)):
            raise ValueError("cylinder arrays must have the same length")
        return self._add(self._fn("brt_scene_add_cylinders"), self._p(c), self._p(a), self._p(b), self._p(h),
                         self._p(R), len(a))

    def commit(self, device: bool = False, stream=None) -> bool:
        """Builds the BVHs (blazert/scene.h:91-115). device=True builds them on the current GPU
        (brt_scene_commit_device: the same trees and records, bit for bit) and leaves the scene uploaded."""
        if device:
            import torch
            st = torch.cuda.current_stream() if stream is None else stream
            check(lib().brt_scene_commit_device(self._h, C.c_void_p(st.cuda_stream)))
        else:
            check(lib().brt_scene_commit(self._h))
        self.has_been_committed = True
        return True

    def read_device_array(self, geom_type: int, which: int) -> np.ndarray:
        """Raw bytes of one device array (capi.ARRAY_*) of a committed geometry type."""
        nbytes = C.c_uint64(0)
        check(lib().brt_scene_read_device_array(self._h, geom_type, which, None, 0, C.byref(nbytes)))
        out = np.zeros(nbytes.value, dtype=np.uint8)
        if nbytes.value:
            check(lib().brt_scene_read_device_array(self._h, geom_type, which, out.ctypes.data_as(C.c_void_p),
                                                    nbytes.value, C.byref(nbytes)))
        return out

    def upload(self) -> None:
        check(lib().brt_scene_upload(self._h))

    def build_stats(self, geom_type: int) -> dict:
        s = capi.BuildStats()
        check(lib().brt_scene_build_stats(self._h, geom_type, C.byref(s)))
        return {n: getattr(s, n) for n, _ in capi.BuildStats._fields_}

    def export_tree(self, geom_type: int):
        """Reference-form tree (pre-order nodes) and the primitive order of the leaves."""
        n = C.c_uint32(0)
        check(lib().brt_scene_export_tree(self._h, geom_type, None, 0, None, C.byref(n)))
        dt = np.dtype([("min", np.float64, 3), ("max", np.float64, 3), ("leaf", np.uint32), ("axis", np.uint32),
                       ("children", np.uint32, 2), ("prim_begin", np.uint32), ("prim_count", np.uint32)])
        nodes = np.zeros(n.value, dtype=dt)
        nprim = self.build_stats(geom_type)["num_primitives"]
        order = np.zeros(max(nprim, 1), dtype=np.uint32)
        check(lib().brt_scene_export_tree(self._h, geom_type, nodes.ctypes.data_as(C.c_void_p), n.value, _up(order),
                                          C.byref(n)))
        return nodes, order[:nprim]

    # -------------------------------------------------------------------------------------------- batched
    def intersect(self, org, dirs, tmin=None, tmax=None, *, precision: int = capi.PRECISION_FP64,
                  any_hit: bool = False, normals: bool = False, hits: dict | None = None, block_size: int = 0,
                  schedule: int = capi.SCHEDULE_PERSISTENT, refill: int = 0, sort: int = capi.SORT_NONE,
                  stack_ring: int = 0, stream=None) -> dict:
        """Batched intersect1 over n rays. org / dirs: (n,3) CUDA tensors or tuples of three columns, in the
        scene type (float64, or float32 for a Scene<float>).

        Returns a dict of CUDA tensors t, u, v (scene type), prim_id, geom_id (int32 holding uint32 bits) and,
        with normals=True, normal (n,3). Asynchronous on `stream` (torch's current stream by default).
        With ``hits`` given (same dict layout), those tensors are used as in/out RayHits.
        """
        import torch

        tdt = torch.float32 if self.f32 else torch.float64
        cols_o = _columns(org, tdt)
        cols_d = _columns(dirs, tdt)
        n = cols_o[0].numel()
        dev = cols_o[0].device
        tmn = None if tmin is None else _scalar_col(tmin, n, dev, tdt)
        tmx = None if tmax is None else _scalar_col(tmax, n, dev, tdt)
        flags = capi.FLAG_ANY_HIT if any_hit else 0
        if hits is None:
            hits = {k: torch.empty(n, dtype=tdt, device=dev) for k in ("t", "u", "v")}
            hits["prim_id"] = torch.empty(n, dtype=torch.int32, device=dev)
            hits["geom_id"] = torch.empty(n, dtype=torch.int32, device=dev)
            if normals:
                hits["normal"] = torch.empty((3, n), dtype=tdt, device=dev)
        else:
            flags |= capi.FLAG_HITS_INOUT
        nrm = hits.get("normal")
        rays = capi.Rays(*[c.data_ptr() for c in cols_o], *[c.data_ptr() for c in cols_d],
                         tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
        H = capi.Hits(hits["t"].data_ptr(), hits["u"].data_ptr(), hits["v"].data_ptr(), hits["prim_id"].data_ptr(),
                      hits["geom_id"].data_ptr(), *([nrm[k].data_ptr() for k in range(3)] if nrm is not None
                                                    else [None, None, None]))
        L = capi.Launch(precision, flags, block_size, schedule, refill, sort, stack_ring)
        st = stream if stream is not None else torch.cuda.current_stream(dev)
        keep = (cols_o, cols_d, tmn, tmx)  # noqa: F841 (inputs alive until the launch is enqueued)
        check(self._fn("brt_intersect")(self._h, C.byref(rays), C.byref(H), n, C.byref(L),
                                        C.c_void_p(st.cuda_stream)))
        return hits

    # -------------------------------------------------------------------------------------------- distances
    def distance_to_surface(self, geom_type: int, points, prim_ids, stream=None):
        """distance_to_surface(primitive prim_ids[i], point i) for device points ((n,3) or three columns) and int32
        ids of one geometry type; FP64 tensor, NaN for ids out of range (brt_distance_to_surface)."""
        import torch
        cols = _columns(points)
        n = cols[0].numel()
        ids = torch.as_tensor(prim_ids).to(device=cols[0].device, dtype=torch.int32).contiguous()
        if ids.numel() != n:
            raise ValueError("one primitive id per point")
        out = torch.empty(n, dtype=torch.float64, device=cols[0].device)
        st = stream if stream is not None else torch.cuda.current_stream(cols[0].device)
        pts = capi.Points(*[c.data_ptr() for c in cols])
        check(lib().brt_distance_to_surface(self._h, geom_type, C.byref(pts), C.c_void_p(ids.data_ptr()), n,
                                            C.c_void_p(out.data_ptr()), C.c_void_p(st.cuda_stream)))
        return out

    def closest_surface(self, points, stream=None) -> dict:
        """Closest surface of the scene for each device point: distance, prim_id, geom_id (brt_closest_surface)."""
        import torch
        cols = _columns(points)
        n = cols[0].numel()
        dev = cols[0].device
        out = {"distance": torch.empty(n, dtype=torch.float64, device=dev),
               "prim_id": torch.empty(n, dtype=torch.int32, device=dev),
               "geom_id": torch.empty(n, dtype=torch.int32, device=dev)}
        st = stream if stream is not None else torch.cuda.current_stream(dev)
        pts = capi.Points(*[c.data_ptr() for c in cols])
        check(lib().brt_closest_surface(self._h, C.byref(pts), n, *[C.c_void_p(out[k].data_ptr())
                                                                   for k in ("distance", "prim_id", "geom_id")],
                                        C.c_void_p(st.cuda_stream)))
        return out

    def traverse(self, geom_type: int, org, dirs, hits: dict, tmin=None, tmax=None, **kw) -> dict:
        """Low-level traverse of one geometry type (blazert/bvh/accel.h:60-120) over device rays: rays that hit get
        their record in `hits` (same layout as intersect's) overwritten, rays that miss keep it."""
        return _traverse(self, geom_type, org, dirs, hits, tmin, tmax, **kw)


def _traverse(scene, geom_type, org, dirs, hits, tmin=None, tmax=None, *, precision=capi.PRECISION_FP64,
              any_hit=False, stream=None):
    import torch

    cols_o, cols_d = _columns(org), _columns(dirs)
    n = cols_o[0].numel()
    dev = cols_o[0].device
    tmn = None if tmin is None else _scalar_col(tmin, n, dev)
    tmx = None if tmax is None else _scalar_col(tmax, n, dev)
    rays = capi.Rays(*[c.data_ptr() for c in cols_o], *[c.data_ptr() for c in cols_d],
                     tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
    L = capi.Launch(precision, capi.FLAG_ANY_HIT if any_hit else 0, 0, capi.SCHEDULE_PERSISTENT, 0, capi.SORT_NONE)
    st = stream if stream is not None else torch.cuda.current_stream(dev)
    keep = (cols_o, cols_d, tmn, tmx)  # noqa: F841
    check(lib().brt_traverse(scene.handle, geom_type, C.byref(rays), C.byref(_hits_struct(hits)), n, C.byref(L),
                             C.c_void_p(st.cuda_stream)))
    return hits


def _alloc_hits(n, dev, normals):
    import torch
    hits = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in ("t", "u", "v")}
    hits["prim_id"] = torch.empty(n, dtype=torch.int32, device=dev)
    hits["geom_id"] = torch.empty(n, dtype=torch.int32, device=dev)
    if normals:
        hits["normal"] = torch.empty((3, n), dtype=torch.float64, device=dev)
    return hits


def _hits_struct(hits):
    nrm = hits.get("normal")
    return capi.Hits(hits["t"].data_ptr(), hits["u"].data_ptr(), hits["v"].data_ptr(), hits["prim_id"].data_ptr(),
                     hits["geom_id"].data_ptr(), *([nrm[k].data_ptr() for k in range(3)] if nrm is not None
                                                   else [None, None, None]))


class Batch:
    """Many (scene, rays) items traced together: one persistent launch per geometry type over all items
    (brt_batch_* in the C ABI). ``items`` is a list of (Scene, origins, directions[, tmin, tmax]); origins and
    directions as for ``Scene.intersect``. ``hits[i]`` holds item i's results after a launch or replay."""

    def __init__(self, items, *, precision: int = capi.PRECISION_FP64, any_hit: bool = False,
                 normals: bool = False, block_size: int = 0, refill: int = 0, stack_ring: int = 0):
        self._keep = []
        self.scenes = []
        self.hits = []
        arr = (capi.BatchItem * max(1, len(items)))()
        for i, it in enumerate(items):
            scene, org, dirs = it[0], it[1], it[2]
            co, cd = _columns(org), _columns(dirs)
            n = co[0].numel()
            dev = co[0].device
            tmn = _scalar_col(it[3], n, dev) if len(it) > 3 and it[3] is not None else None
            tmx = _scalar_col(it[4], n, dev) if len(it) > 4 and it[4] is not None else None
            h = _alloc_hits(n, dev, normals)
            self._keep.extend([co, cd, tmn, tmx])
            self.scenes.append(scene)
            self.hits.append(h)
            arr[i].scene = scene.handle
            arr[i].rays = capi.Rays(*[c.data_ptr() for c in co], *[c.data_ptr() for c in cd],
                                    tmn.data_ptr() if tmn is not None else None,
                                    tmx.data_ptr() if tmx is not None else None)
            arr[i].hits = _hits_struct(h)
            arr[i].n = n
        L = capi.Launch(precision, capi.FLAG_ANY_HIT if any_hit else 0, block_size, capi.SCHEDULE_PERSISTENT, refill,
                        capi.SORT_NONE, stack_ring)
        h = C.c_void_p()
        check(lib().brt_batch_create(arr, len(items), C.byref(L), C.byref(h)))
        self._h = h

    def _stream(self, stream):
        import torch
        return stream if stream is not None else torch.cuda.current_stream()

    def launch(self, stream=None):
        check(lib().brt_batch_launch(self._h, C.c_void_p(self._stream(stream).cuda_stream)))

    def capture(self, stream):
        """Records the launch sequence on `stream` (not the default stream) into a hipGraph."""
        check(lib().brt_batch_capture(self._h, C.c_void_p(stream.cuda_stream)))

    def replay(self, stream=None):
        check(lib().brt_batch_replay(self._h, C.c_void_p(self._stream(stream).cuda_stream)))

    @property
    def handle(self):
        return self._h

    def close(self):
        h = getattr(self, "_h", None)
        self._h = None
        if h:
            try:
                lib().brt_batch_destroy(h)
            except Exception:
                pass

    def __del__(self):
        self.close()


def _columns(x, dtype=None):
    """(n,3) tensor or three columns -> three contiguous columns (float64 unless `dtype`) on the same device."""
    import torch

    dtype = dtype or torch.float64
    if isinstance(x, (list, tuple)) and len(x) == 3:
        return [torch.as_tensor(c, dtype=dtype).contiguous() for c in x]
    x = torch.as_tensor(x, dtype=dtype)
    if not (x.dim() == 2 and x.shape[1] == 3):
        raise ValueError("expected an (n,3) tensor or a tuple of three columns")
    x = x.t()
    return [x[k].contiguous() for k in range(3)]


def _scalar_col(v, n, dev, dtype=None):
    import torch

    t = torch.as_tensor(v, dtype=dtype or torch.float64, device=dev)
    if t.dim() == 0:
        t = t.expand(n)
    return t.contiguous()


# ------------------------------------------------------------------------------------------------ drop-in
@dataclass
class Ray:
    """blazert::Ray<double> (blazert/ray.h:20-66): origin, direction, min/max hit distance, any_hit."""
    origin: tuple
    direction: tuple
    min_hit_distance: float = 0.0
    max_hit_distance: float = DBL_MAX
    any_hit: bool = False


@dataclass
class RayHit:
    """blazert::RayHit<double> (blazert/ray.h:72-85) with the reference defaults."""
    normal: list = field(default_factory=lambda: [0.0, 0.0, 0.0])
    uv: list = field(default_factory=lambda: [0.0, 0.0])
    hit_distance: float = DBL_MAX
    prim_id: int = INVALID_ID
    geom_id: int = INVALID_ID


def intersect1(scene: Scene, ray: Ray, rayhit: RayHit) -> bool:
    """intersect1(scene, ray, rayhit) (blazert/scene.h:161-211): one ray on the GPU, rayhit updated in place.
    For a Scene<float> every value is rounded to float32 first, as Ray<float> / RayHit<float> hold them."""
    f32 = scene.f32
    ct = C.c_float if f32 else C.c_double
    mx = float(np.finfo(np.float32).max) if f32 else DBL_MAX
    H = capi.Hit1F if f32 else capi.Hit1
    h = H((ct * 3)(*rayhit.normal), rayhit.uv[0], rayhit.uv[1], min(rayhit.hit_distance, mx), rayhit.prim_id,
          rayhit.geom_id)
    o = np.asarray(ray.origin, dtype=scene.dtype)
    d = np.asarray(ray.direction, dtype=scene.dtype)
    res = C.c_int(0)
    fn = lib().brt_intersect1_f32 if f32 else lib().brt_intersect1
    check(fn(scene.handle, o.ctypes.data_as(C.POINTER(ct)), d.ctypes.data_as(C.POINTER(ct)),
             float(ray.min_hit_distance), min(float(ray.max_hit_distance), mx),
             capi.FLAG_ANY_HIT if ray.any_hit else 0, C.byref(h), C.byref(res)))
    rayhit.normal = [h.normal[0], h.normal[1], h.normal[2]]
    rayhit.uv = [h.u, h.v]
    rayhit.hit_distance = h.t
    rayhit.prim_id = h.prim_id
    rayhit.geom_id = h.geom_id
    return bool(res.value)
