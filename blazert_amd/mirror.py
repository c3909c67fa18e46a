"""Differentiable multi-bounce mirror tracing on a triangle mesh (brt_mirror_trace / brt_mirror_adjoint).

No counterpart in blazeRT: this extends its ``intersect1`` (blazert/scene.h:161-211) with a mirror bounce. Bounce 0
traces the given rays; bounce b >= 1 starts at bounce b-1's hit position and follows the mirror image of the unit
direction about the unit face normal ``normalize(cross(v1 - v0, v2 - v0))`` (blazert/primitives/trimesh.h:125-129).
Each bounce's hit is the exact FP64 closest hit. ``MirrorTrace`` wraps the pair as a torch.autograd.Function whose
backward is the HIP adjoint kernel: it re-derives every bounce from the stored face ids and scatter-adds dL/dvertex
in FP64.
"""
from __future__ import annotations

import ctypes as C

from . import capi
from .capi import check, lib
from .scene import Scene, _columns, _OnStream, _scalar_col

DEFAULT_BOUNCE_TMIN = 1e-9


def _rays(org, dirs):
    co = _columns(org, what="origins")
    dev = co[0].device
    cd = _columns(dirs, dev=dev, what="directions")
    if cd[0].numel() != co[0].numel():
        raise ValueError(f"directions: expected {co[0].numel()} rays, got {cd[0].numel()}")
    return co, cd, dev


def _cuda_on(x, dev, what):
    import torch
    if not torch.is_tensor(x) or not x.is_cuda or x.device != dev:
        raise ValueError(f"{what}: expected a CUDA tensor on {dev}")


def mirror_trace(scene: Scene, org, dirs, bounces: int = 3, bounce_tmin: float = DEFAULT_BOUNCE_TMIN, tmin=None,
                 tmax=None, *, uv: bool = False, block_size: int = 0, schedule: int = capi.SCHEDULE_PERSISTENT,
                 refill: int = 0, precision: int = capi.PRECISION_FP64, stream=None) -> dict:
    """Traces `bounces` mirror bounces of n rays. org / dirs: (n,3) float64 CUDA tensors or three columns.

    Returns CUDA tensors t (bounces, n), prim_id (bounces, n; int32 holding uint32 bits, -1 = miss), position and
    direction (3, bounces, n) (x, y, z planes: the next bounce's origin and direction; NaN after a miss) and, with
    uv=True, u and v (bounces, n). precision=PRECISION_FP32_TRAV_FP64_HIT traces each bounce with FP32 traversal
    and FP64 hits. Asynchronous on `stream`.
    """
    import torch

    co, cd, dev = _rays(org, dirs)
    n = co[0].numel()
    tmn = None if tmin is None else _scalar_col(tmin, n, dev, what="tmin")
    tmx = None if tmax is None else _scalar_col(tmax, n, dev, what="tmax")
    B = int(bounces)
    out = {"t": torch.empty((B, n), dtype=torch.float64, device=dev),
           "prim_id": torch.empty((B, n), dtype=torch.int32, device=dev),
           "position": torch.empty((3, B, n), dtype=torch.float64, device=dev),
           "direction": torch.empty((3, B, n), dtype=torch.float64, device=dev)}
    if uv:
        out["u"] = torch.empty((B, n), dtype=torch.float64, device=dev)
        out["v"] = torch.empty((B, n), dtype=torch.float64, device=dev)
    rays = capi.Rays(*[c.data_ptr() for c in co], *[c.data_ptr() for c in cd],
                     tmn.data_ptr() if tmn is not None else None, tmx.data_ptr() if tmx is not None else None)
    P, D = out["position"], out["direction"]
    mo = capi.MirrorOut(out["t"].data_ptr(), out["u"].data_ptr() if uv else None,
                        out["v"].data_ptr() if uv else None, out["prim_id"].data_ptr(),
                        P[0].data_ptr(), P[1].data_ptr(), P[2].data_ptr(), D[0].data_ptr(), D[1].data_ptr(),
                        D[2].data_ptr())
    L = capi.Launch(precision, 0, block_size, schedule, refill, capi.SORT_NONE)
    on = _OnStream(stream, dev)
    check(lib().brt_mirror_trace(scene.handle, C.byref(rays), n, B, float(bounce_tmin), C.byref(L), C.byref(mo),
                                 on.handle))
    on.hold(co, cd, tmn, tmx, out)
    return out


def mirror_adjoint(scene: Scene, org, dirs, prim_id, grad_t=None, grad_position=None, *, grad_vertices=None,
                   want_rays: bool = True, stream=None):
    """dL/dvertex (nv,3), dL/dorigin (n,3) and dL/ddirection (n,3) for upstream dL/dt (bounces, n) and
    dL/dposition (3, bounces, n) of mirror_trace's outputs. `grad_vertices` (nv,3 float64, contiguous) is
    accumulated into when given."""
    import torch

    co, cd, dev = _rays(org, dirs)
    n = co[0].numel()
    _cuda_on(prim_id, dev, "prim_id")
    for name, g in (("grad_t", grad_t), ("grad_position", grad_position), ("grad_vertices", grad_vertices)):
        if g is not None:
            _cuda_on(g, dev, name)
    prim = prim_id.contiguous()
    if prim.dim() != 2 or prim.shape[1] != n or prim.dtype not in (torch.int32, torch.uint32):
        raise ValueError("prim_id must be a (bounces, n) int32 tensor from mirror_trace")
    B = prim.shape[0]
    if grad_t is not None and tuple(grad_t.shape) != (B, n):
        raise ValueError("grad_t must have shape (bounces, n)")
    if grad_position is not None and tuple(grad_position.shape) != (3, B, n):
        raise ValueError("grad_position must have shape (3, bounces, n)")
    gt = None if grad_t is None else grad_t.to(torch.float64).contiguous()
    gp = None if grad_position is None else grad_position.to(torch.float64).contiguous()
    if grad_vertices is None:
        grad_vertices = torch.zeros((scene.num_vertices, 3), dtype=torch.float64, device=dev)
    elif (tuple(grad_vertices.shape) != (scene.num_vertices, 3) or grad_vertices.dtype != torch.float64
          or not grad_vertices.is_contiguous()):
        raise ValueError("grad_vertices must be a contiguous (nv, 3) float64 tensor")
    go = torch.empty((3, n), dtype=torch.float64, device=dev) if want_rays else None
    gd = torch.empty((3, n), dtype=torch.float64, device=dev) if want_rays else None
    g = capi.MirrorGrads(gt.data_ptr() if gt is not None else None,
                         *([gp[k].data_ptr() for k in range(3)] if gp is not None else [None] * 3),
                         grad_vertices.data_ptr(),
                         *([go[k].data_ptr() for k in range(3)] if want_rays else [None] * 3),
                         *([gd[k].data_ptr() for k in range(3)] if want_rays else [None] * 3))
    rays = capi.Rays(*[c.data_ptr() for c in co], *[c.data_ptr() for c in cd], None, None)
    on = _OnStream(stream, dev)
    check(lib().brt_mirror_adjoint(scene.handle, C.byref(rays), n, B, prim.data_ptr(), C.byref(g), on.handle))
    on.hold(co, cd, prim, gt, gp, grad_vertices, go, gd)
    return grad_vertices, (go.t() if want_rays else None), (gd.t() if want_rays else None)


def _make_function():
    import torch

    class MirrorTrace(torch.autograd.Function):
        """(vertices, origins, directions) -> (t (bounces, n), positions (bounces, n, 3)).

        `vertices` must be the (nv,3) float64 vertices the scene was committed with: the forward pass traces the
        committed scene and the gradient is taken there. Rays that miss contribute nothing from that bounce on.
        """

        @staticmethod
        def forward(ctx, vertices, origins, directions, scene, bounces, bounce_tmin):
            if vertices.shape != (scene.num_vertices, 3):
                raise ValueError("vertices must be the committed mesh vertices, shape (nv, 3)")
            out = mirror_trace(scene, origins.detach(), directions.detach(), bounces, bounce_tmin)
            ctx.scene = scene
            ctx.save_for_backward(origins.detach(), directions.detach(), out["prim_id"])
            t = out["t"]
            pos = out["position"].permute(1, 2, 0)
            ctx.mark_non_differentiable(out["prim_id"])
            return t, pos, out["prim_id"]

        @staticmethod
        def backward(ctx, grad_t, grad_pos, _grad_prim):
            origins, directions, prim = ctx.saved_tensors
            gp = None if grad_pos is None else grad_pos.permute(2, 0, 1)
            gv, go, gd = mirror_adjoint(ctx.scene, origins, directions, prim, grad_t, gp)
            return gv, go, gd, None, None, None

    return MirrorTrace


_fn = None


def mirror_trace_autograd(scene: Scene, vertices, origins, directions, bounces: int = 3,
                          bounce_tmin: float = DEFAULT_BOUNCE_TMIN):
    """Differentiable mirror tracing: returns (t, positions, prim_id); t and positions carry gradients to
    vertices, origins and directions. Misses have t = DBL_MAX / NaN positions: mask them (prim_id == -1) before
    building a loss, their upstream gradient is ignored."""
    global _fn
    if _fn is None:
        _fn = _make_function()
    return _fn.apply(vertices, origins, directions, scene, int(bounces), float(bounce_tmin))
