"""Autotuning of launch parameters and BVH leaf size, with a per-device JSON cache.

The measurements are made by the library (brt_autotune_launch, brt_batch_autotune, brt_scene_tune_leaf in
include/blazert_amd.h): candidates are timed with HIP events on the caller's own device rays and every candidate
computes the same hits, so a tuned launch is only faster. This module keeps the choices in a JSON file keyed by
device, precision, flags, the scene's primitive counts and the ray count (rounded to a power of two), so a workload
is tuned once per device. Cache file: $BLAZERT_AMD_TUNE_CACHE, else ~/.cache/blazert_amd/tune.json.
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os

from . import capi
from .capi import check, lib
from .scene import Scene, _alloc_hits, _columns, _hits_struct, _OnStream

LAUNCH_KEYS = ("block_size", "refill", "stack_ring", "sort", "pool_size")


def cache_path() -> str:
    return os.environ.get("BLAZERT_AMD_TUNE_CACHE") or os.path.join(os.path.expanduser("~"), ".cache", "blazert_amd",
                                                                   "tune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    data = _load()
    data[key] = value
    path = cache_path()
    os.makedirs(os.path.dirname(path), exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def _device() -> str:
    import torch
    return torch.cuda.get_device_name()


def scene_signature(scene: Scene) -> str:
    parts = []
    for t in range(4):
        try:
            st = scene.build_stats(t)
        except capi.BlazertError:
            continue
        parts.append(f"{t}:{st['num_primitives']}:{st['max_tree_depth']}")
    return ",".join(parts) or "empty"


def _bucket(n: int) -> int:
    return 1 << max(0, int(math.log2(max(n, 1)) + 0.5))


def _result(r: capi.TuneResult) -> dict:
    L = r.launch
    return {"block_size": L.block_size, "refill": L.refill, "stack_ring": L.stack_ring, "sort": L.sort,
            "pool_size": L.pool_size, "min_leaf_primitives": r.min_leaf_primitives, "candidates": r.candidates,
            "best_ms": r.best_ms, "default_ms": r.default_ms}


def _rays(org, dirs):
    co = _columns(org, what="origins")
    cd = _columns(dirs, dev=co[0].device, what="directions")
    if cd[0].numel() != co[0].numel():
        raise ValueError(f"directions: expected {co[0].numel()} rays, got {cd[0].numel()}")
    return co, cd, capi.Rays(*[c.data_ptr() for c in co], *[c.data_ptr() for c in cd], None, None)


def tune_launch(scene: Scene, org, dirs, *, precision: int = capi.PRECISION_FP64, any_hit: bool = False,
                repeats: int = 3, use_cache: bool = True, stream=None) -> dict:
    """Fastest launch parameters for this scene and ray set (org / dirs: device (n,3) tensors or columns, a
    representative sample of the workload). Returns a dict whose LAUNCH_KEYS entries are Scene.intersect keywords."""
    co, cd, rays = _rays(org, dirs)
    n = co[0].numel()
    flags = capi.FLAG_ANY_HIT if any_hit else 0
    key = f"launch|{_device()}|p{precision}|f{flags}|{scene_signature(scene)}|n{_bucket(n)}"
    if use_cache:
        hit = _load().get(key)
        if hit is not None and all(k in hit for k in LAUNCH_KEYS):  # entries from older versions are re-tuned
            return hit
    hits = _alloc_hits(n, co[0].device, False)
    on = _OnStream(stream, co[0].device)
    r = capi.TuneResult()
    check(lib().brt_autotune_launch(scene.handle, C.byref(rays), C.byref(_hits_struct(hits)), n, precision, flags,
                                    repeats, C.byref(r), on.handle))
    on.hold(co, cd, hits)
    out = _result(r)
    if use_cache:
        _store(key, out)
    return out


def tune_batch(batch, *, repeats: int = 3, stream=None) -> dict:
    """Tunes a Batch in place (it keeps the fastest block size / refill / stack ring / pool size); call before
    capture."""
    import torch

    st = stream if stream is not None else torch.cuda.current_stream()
    r = capi.TuneResult()
    check(lib().brt_batch_autotune(batch.handle, repeats, C.byref(r), C.c_void_p(st.cuda_stream)))
    return _result(r)


def tune_leaf(scene: Scene, org, dirs, *, precision: int = capi.PRECISION_FP64, leaf_sizes=(1, 2, 4, 8),
              repeats: int = 3, stream=None) -> dict:
    """Rebuilds the scene with each leaf size, times it on the rays and leaves it built with the fastest."""
    import numpy as np

    co, cd, rays = _rays(org, dirs)
    n = co[0].numel()
    hits = _alloc_hits(n, co[0].device, False)
    sizes = np.ascontiguousarray(leaf_sizes, dtype=np.uint32)
    on = _OnStream(stream, co[0].device)
    r = capi.TuneResult()
    check(lib().brt_scene_tune_leaf(scene.handle, C.byref(rays), C.byref(_hits_struct(hits)), n, precision,
                                    sizes.ctypes.data_as(C.POINTER(C.c_uint32)), len(sizes), repeats, C.byref(r),
                                    on.handle))
    on.hold(co, cd, hits)
    scene.has_been_committed = True
    return _result(r)
