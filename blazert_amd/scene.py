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
                  stack_ring: int = 0, pool_size: int = 0, stream=None) -> dict:
        """Batched intersect1 over n rays. org / dirs: (n,3) CUDA tensors or tuples of three columns, in the
        scene type (float64, or float32 for a Scene<float>).

        Returns a dict of CUDA tensors t, u, v (scene type), prim_id, geom_id (int32 holding uint32 bits) and,
        with normals=True, normal (n,3). Asynchronous on `stream` (torch's current stream by default).
        With ``hits`` given (same dict layout), those tensors are used as in/out RayHits.
        """
        import torch

        tdt = torch.float32 if self.f32 else torch.float64
        cols_o = _columns(org, tdt, what="origins")
        dev = cols_o[0].device
        cols_d = _columns(dirs, tdt, dev, what="directions")
        n = cols_o[0].numel()
        if cols_d[0].numel() != n:
            raise ValueError(f"directions: expected {n} rays, got {cols_d[0].numel()}")
        tmn = None if tmin is None else _scalar_col(tmin, n, dev, tdt, "tmin")
        tmx = None if tmax is None else _scalar_col(tmax, n, dev, tdt, "tmax")
        flags = capi.FLAG_ANY_HIT if any_hit else 0
        if hits is None:
            hits = {k: torch.empty(n, dtype=tdt, device=dev) for k in ("t", "u", "v")}
            hits["prim_id"] = torch.empty(n, dtype=torch.int32, device=dev)
            hits["geom_id"] = torch.empty(n, dtype=torch.int32, device=dev)
            if normals:
                hits["normal"] = torch.empty((3, n), dtype=tdt, device=dev)
        else:
            _check_hits(hits, n, dev, tdt)
            flags |= capi.FLAG_HITS_INOUT
        nrm = hits.get("normal")
        rays = capi.Rays(*[c.data_ptr() for c in cols_o], *[c.data_ptr() for c in cols_d],
                         tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
        H = capi.Hits(hits["t"].data_ptr(), hits["u"].data_ptr(), hits["v"].data_ptr(), hits["prim_id"].data_ptr(),
                      hits["geom_id"].data_ptr(), *([nrm[k].data_ptr() for k in range(3)] if nrm is not None
                                                    else [None, None, None]))
        L = capi.Launch(precision, flags, block_size, schedule, refill, sort, stack_ring, pool_size)
        on = _OnStream(stream, dev)
        check(self._fn("brt_intersect")(self._h, C.byref(rays), C.byref(H), n, C.byref(L), on.handle))
        on.hold(cols_o, cols_d, tmn, tmx, hits)
        return hits

    # -------------------------------------------------------------------------------------------- distances
    def distance_to_surface(self, geom_type: int, points, prim_ids, stream=None):
        """distance_to_surface(primitive prim_ids[i], point i) for device points ((n,3) or three columns) and int32
        ids of one geometry type; FP64 tensor, NaN for ids out of range (brt_distance_to_surface)."""
        import torch
        cols = _columns(points, what="points")
        n = cols[0].numel()
        dev = cols[0].device
        if not torch.is_tensor(prim_ids) or not prim_ids.is_cuda or prim_ids.device != dev:
            raise ValueError(f"prim_ids: expected a CUDA tensor on {dev}")
        if prim_ids.dtype not in (torch.int32, torch.int64, torch.uint32) or prim_ids.numel() != n:
            raise ValueError(f"prim_ids: expected {n} integer ids")
        ids = prim_ids.to(torch.int32).reshape(-1).contiguous()
        out = torch.empty(n, dtype=torch.float64, device=dev)
        on = _OnStream(stream, dev)
        pts = capi.Points(*[c.data_ptr() for c in cols])
        check(lib().brt_distance_to_surface(self._h, geom_type, C.byref(pts), C.c_void_p(ids.data_ptr()), n,
                                            C.c_void_p(out.data_ptr()), on.handle))
        on.hold(cols, ids, out)
        return out

    def closest_surface(self, points, stream=None) -> dict:
        """Closest surface of the scene for each device point: distance, prim_id, geom_id (brt_closest_surface)."""
        import torch
        cols = _columns(points, what="points")
        n = cols[0].numel()
        dev = cols[0].device
        out = {"distance": torch.empty(n, dtype=torch.float64, device=dev),
               "prim_id": torch.empty(n, dtype=torch.int32, device=dev),
               "geom_id": torch.empty(n, dtype=torch.int32, device=dev)}
        on = _OnStream(stream, dev)
        pts = capi.Points(*[c.data_ptr() for c in cols])
        check(lib().brt_closest_surface(self._h, C.byref(pts), n, *[C.c_void_p(out[k].data_ptr())
                                                                   for k in ("distance", "prim_id", "geom_id")],
                                        on.handle))
        on.hold(cols, out)
        return out

    def traverse(self, geom_type: int, org, dirs, hits: dict, tmin=None, tmax=None, **kw) -> dict:
        """Low-level traverse of one geometry type (blazert/bvh/accel.h:60-120) over device rays: rays that hit get
        their record in `hits` (same layout as intersect's) overwritten, rays that miss keep it."""
        return _traverse(self, geom_type, org, dirs, hits, tmin, tmax, **kw)


def _traverse(scene, geom_type, org, dirs, hits, tmin=None, tmax=None, *, precision=capi.PRECISION_FP64,
              any_hit=False, stream=None):
    import torch

    if scene.f32:
        raise ValueError("traverse is available for Scene<double> only")
    cols_o = _columns(org, what="origins")
    dev = cols_o[0].device
    cols_d = _columns(dirs, dev=dev, what="directions")
    n = cols_o[0].numel()
    if cols_d[0].numel() != n:
        raise ValueError(f"directions: expected {n} rays, got {cols_d[0].numel()}")
    _check_hits(hits, n, dev, torch.float64)
    tmn = None if tmin is None else _scalar_col(tmin, n, dev, what="tmin")
    tmx = None if tmax is None else _scalar_col(tmax, n, dev, what="tmax")
    rays = capi.Rays(*[c.data_ptr() for c in cols_o], *[c.data_ptr() for c in cols_d],
                     tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
    L = capi.Launch(precision, capi.FLAG_ANY_HIT if any_hit else 0, 0, capi.SCHEDULE_PERSISTENT, 0, capi.SORT_NONE)
    on = _OnStream(stream, dev)
    check(lib().brt_traverse(scene.handle, geom_type, C.byref(rays), C.byref(_hits_struct(hits)), n, C.byref(L),
                             on.handle))
    on.hold(cols_o, cols_d, tmn, tmx, hits)
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
                 normals: bool = False, block_size: int = 0, refill: int = 0, stack_ring: int = 0,
                 pool_size: int = 0):
        self._keep = []
        self.scenes = []
        self.hits = []
        arr = (capi.BatchItem * max(1, len(items)))()
        import torch

        devices = set()
        for i, it in enumerate(items):
            scene, org, dirs = it[0], it[1], it[2]
            if scene.f32:
                raise ValueError("batches hold Scene<double> items only")
            co = _columns(org, what=f"item {i} origins")
            dev = co[0].device
            cd = _columns(dirs, dev=dev, what=f"item {i} directions")
            n = co[0].numel()
            if cd[0].numel() != n:
                raise ValueError(f"item {i}: {n} origins but {cd[0].numel()} directions")
            devices.add(dev)
            tmn = _scalar_col(it[3], n, dev, what="tmin") if len(it) > 3 and it[3] is not None else None
            tmx = _scalar_col(it[4], n, dev, what="tmax") if len(it) > 4 and it[4] is not None else None
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
        if len(devices) > 1:
            raise ValueError("all batch items must be on one GPU")
        # the batch's columns and hit tensors live as long as the batch; they were written on the current stream,
        # and launches, captures and replays may run on any stream, so finish those writes here once
        for d in devices:
            torch.cuda.current_stream(d).synchronize()
        L = capi.Launch(precision, capi.FLAG_ANY_HIT if any_hit else 0, block_size, capi.SCHEDULE_PERSISTENT, refill,
                        capi.SORT_NONE, stack_ring, pool_size)
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


def _columns(x, dtype=None, dev=None, what="rays"):
    """(n,3) tensor or three columns -> three contiguous columns of `dtype` (float64 unless given), all on one GPU.

    The columns are handed to kernels as raw device pointers, so anything else is refused before a launch: host
    data (numpy arrays, CPU tensors), columns on different devices or of different lengths, a wrong shape, or a
    device other than `dev` when given. A floating tensor of another precision is converted (a copy on the same
    device); an integer or complex one is refused."""
    import torch

    dtype = dtype or torch.float64
    if isinstance(x, (list, tuple)) and len(x) == 3:
        cols = list(x)
    elif torch.is_tensor(x) and x.dim() == 2 and x.shape[1] == 3:
        cols = [x[:, k] for k in range(3)]
    else:
        raise ValueError(f"{what}: expected an (n,3) CUDA tensor or a tuple of three CUDA columns")
    for c in cols:
        if not torch.is_tensor(c):
            raise ValueError(f"{what}: expected CUDA tensors (host arrays cannot be read by the kernels)")
        if not c.is_floating_point():
            raise ValueError(f"{what}: expected floating-point tensors, got {c.dtype}")
        if c.dim() != 1:
            raise ValueError(f"{what}: columns must be 1-D")
    if len({c.numel() for c in cols}) != 1:
        raise ValueError(f"{what}: the three columns differ in length")
    if not all(c.is_cuda for c in cols):
        raise ValueError(f"{what}: expected CUDA tensors (host arrays cannot be read by the kernels)")
    d = cols[0].device if dev is None else torch.device(dev)
    if any(c.device != d for c in cols):
        raise ValueError(f"{what}: all tensors must be on {d}")
    return [c.to(dtype).contiguous() for c in cols]


def _scalar_col(v, n, dev, dtype=None, what="tmin/tmax"):
    """A per-ray FP column of length n on `dev`: a Python scalar is broadcast; a tensor must already be a CUDA tensor
    on `dev` with n elements (a host array cannot be read by the kernels)."""
    import torch

    dtype = dtype or torch.float64
    if not torch.is_tensor(v):
        if np.ndim(v) != 0:
            raise ValueError(f"{what}: give a scalar or a CUDA tensor with one value per ray")
        return torch.full((n,), float(v), dtype=dtype, device=dev)
    if not v.is_cuda or v.device != torch.device(dev):
        raise ValueError(f"{what}: expected a CUDA tensor on {dev}")
    if v.dim() == 0:
        return v.to(dtype).expand(n).contiguous()
    if v.numel() != n:
        raise ValueError(f"{what}: expected {n} values, got {v.numel()}")
    return v.to(dtype).reshape(-1).contiguous()


def _check_hits(hits, n, dev, fdtype):
    """A caller's in/out hit record: CUDA tensors on `dev`, contiguous, t/u/v of the scene type with n elements,
    prim_id/geom_id int32 with n elements, and an optional (3, n) normal. Raises ValueError otherwise; nothing is
    converted, since the kernels write into these tensors."""
    import torch

    spec = {"t": (fdtype, (n,)), "u": (fdtype, (n,)), "v": (fdtype, (n,)), "prim_id": (torch.int32, (n,)),
            "geom_id": (torch.int32, (n,))}
    if hits.get("normal") is not None:
        spec["normal"] = (fdtype, (3, n))
    for k, (dt, shape) in spec.items():
        h = hits.get(k)
        if not torch.is_tensor(h):
            raise ValueError(f"hits[{k!r}]: expected a CUDA tensor on {dev}")
        if h.dtype != dt or tuple(h.shape) != shape or not h.is_contiguous():
            raise ValueError(f"hits[{k!r}]: expected a contiguous {dt} tensor of shape {shape}, got {h.dtype} "
                             f"{tuple(h.shape)}")
    for k in spec:
        if not hits[k].is_cuda or hits[k].device != torch.device(dev):
            raise ValueError(f"hits[{k!r}]: expected a CUDA tensor on {dev}")


class _OnStream:
    """Stream discipline of one asynchronous call. Temporaries (converted columns, broadcast tmin/tmax, new hit
    tensors) are written on torch's current stream; the kernels run on `stream`. So `stream` first waits for the
    current stream, and every tensor the kernels touch is recorded on `stream`, which keeps the caching allocator
    from handing its memory to another tensor before the kernels are done with it."""

    def __init__(self, stream, dev):
        import torch

        cur = torch.cuda.current_stream(dev)
        self.stream = stream if stream is not None else cur
        if self.stream != cur:
            self.stream.wait_stream(cur)

    @property
    def handle(self):
        return C.c_void_p(self.stream.cuda_stream)

    def hold(self, *objs):
        import torch

        for o in objs:
            if o is None:
                continue
            if isinstance(o, dict):
                self.hold(*o.values())
            elif isinstance(o, (list, tuple)):
                self.hold(*o)
            elif torch.is_tensor(o) and o.is_cuda:
                o.record_stream(self.stream)


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
