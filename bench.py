"""Closest-hit throughput on one MI355X: 1,310,720-triangle displaced icosphere, 4,194,304 incoherent rays.

A step is one batched intersect1 over all rays already resident in HBM (FP64 drop-in path). The same rays are
also traced with FP32 traversal + FP64 hit, and the 1920x1080 coherent set is timed beside them. Prints one
JSON line (the value is FP64 Mrays/s). Algorithmic bytes per ray come from the oracle's replay of the
reference traversal on the first rays of the set (node visits x 56 B + triangle tests x 76 B + 96 B ray/hit).
The CPU baseline is the oracle (the reference-exact C restatement) timed on this host on a bounded sample.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "closest-hit Mrays/s: FP64 vs FP32-trav+FP64-hit @1.3M tris; adjoint Mrays/s"
HBM_PEAK_GBS = 8000.0
S_NODE64, S_NODE32, S_TRI, S_RAY_HIT = 56, 32, 76, 96


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launch(fn, steps, warmup):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ms = []
    t0 = time.perf_counter()
    for _ in range(steps):
        ev0.record()
        fn()
        ev1.record()
        ev1.synchronize()
        ms.append(ev0.elapsed_time(ev1))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return float(np.mean(ms)), wall


def classify_disagreements(osc, org, dirs, idx, other=None, cyl_geom=None) -> dict:
    """Counts of the split/FP64 disagreements by class (tie, cylinder, reference overflow, unexplained), from the
    oracle's classifier. `other` holds the split path's t and geom_id on the disagreeing rays."""
    from oracle import oracle as OR
    ot, og = other if other is not None else (np.full(len(idx), np.nan), np.full(len(idx), 0xFFFFFFFF, np.uint32))
    c = OR.classify_split_disagreements(osc, org, dirs, idx, ot, og, cyl_geom=cyl_geom)
    c.pop("unexplained_idx")
    return c


def split_agreement(h64, h32) -> tuple:
    same = (h64["prim_id"] == h32["prim_id"]) & (h64["geom_id"] == h32["geom_id"])
    hit = same & (h64["geom_id"] != -1)
    rel = ((h64["t"] - h32["t"]).abs() / h64["t"].abs())[hit]
    max_rel = float(rel.max().item()) if rel.numel() else 0.0
    idx = (~same).nonzero().flatten()
    other = (h32["t"][idx].cpu().numpy(), h32["geom_id"][idx].cpu().numpy().view(np.uint32))
    return float(same.double().mean().item()), max_rel, idx.cpu().numpy(), other


def mixed_leg(args, dev) -> dict:
    """Mixed-primitive scene (200k triangles as displaced icosphere patches merged into one mesh, 50k spheres, 10k
    cylinders, 100 planes) x 8,388,608 incoherent rays: FP64 and FP32-traversal/FP64-hit throughput, hit-id
    agreement with every disagreement classified by the oracle, and FP64 bit-exactness on a sample."""
    import torch

    from blazert_amd import PRECISION_FP32_TRAV_FP64_HIT, PRECISION_FP64, Scene, workloads as W
    from oracle import oracle as OR

    m = W.mixed_scene()
    org, dirs = W.mixed_rays()
    n = len(org)
    scene = Scene()
    osc = OR.Scene()
    for s in (scene, osc):
        s.add_mesh(*m["mesh"])
        s.add_spheres(*m["spheres"])
        cyl_geom = s.add_cylinders(*m["cylinders"])
        s.add_planes(*m["planes"])
    t0 = time.perf_counter()
    scene.commit()
    scene.upload()
    build_s = time.perf_counter() - t0
    osc.commit()
    O = tuple(torch.from_numpy(np.ascontiguousarray(org[:, k])).to(dev) for k in range(3))
    D = tuple(torch.from_numpy(np.ascontiguousarray(dirs[:, k])).to(dev) for k in range(3))
    steps = max(1, args.steps // 2)
    ms64, _ = time_launch(lambda: scene.intersect(O, D, precision=PRECISION_FP64), steps, args.warmup)
    ms32, _ = time_launch(lambda: scene.intersect(O, D, precision=PRECISION_FP32_TRAV_FP64_HIT), steps, args.warmup)
    h64 = scene.intersect(O, D, precision=PRECISION_FP64)
    h32 = scene.intersect(O, D, precision=PRECISION_FP32_TRAV_FP64_HIT)
    torch.cuda.synchronize()
    agree, max_rel, idx, other = split_agreement(h64, h32)
    log(f"mixed: fp64 {n / ms64 / 1e3:.1f} / fp32-trav {n / ms32 / 1e3:.1f} Mrays/s, agreement {agree:.7f}")
    ns = 1 << 18
    ref = osc.intersect_batch(org[:ns], dirs[:ns], normals=False, trace=True,
                              nthreads=len(os.sched_getaffinity(0)))
    ok = ref["trace"]["overflow"] == 0
    gt = h64["t"][:ns].cpu().numpy()
    gp = h64["prim_id"][:ns].cpu().numpy().view(np.uint32)
    gg = h64["geom_id"][:ns].cpu().numpy().view(np.uint32)
    exact = bool(np.array_equal(gt[ok].view(np.uint64), ref["t"][ok].view(np.uint64))
                 and np.array_equal(gp[ok], ref["prim_id"][ok]) and np.array_equal(gg[ok], ref["geom_id"][ok]))
    return {"workload": f"{len(m['mesh'][1])} tris + 50,000 spheres + 10,000 cylinders + 100 planes x {n} "
                        "incoherent rays (origins U[-22,22]^3)",
            "fp64_mrays_s": round(n / ms64 / 1e3, 3), "fp32_trav_mrays_s": round(n / ms32 / 1e3, 3),
            "speedup": round(ms64 / ms32, 3), "hit_id_agreement": agree, "max_rel_dt_agreeing": max_rel,
            "disagreements": classify_disagreements(osc, org, dirs, idx, other, cyl_geom),
            "commit_s": round(build_s, 3),
            "fp64_bit_exact_vs_oracle_sample": exact, "sample_rays": ns}


def batch_leg(args, dev) -> dict:
    """32 triangle scenes (displaced icospheres and tori, 1k-500k triangles, log-uniform) each with its own
    incoherent rays (16k-2M, log-uniform), split precision. Leaf size is tuned per scene and block size / refill /
    LDS stack depth for the whole batch by the library's autotuner; the batch is captured once in a hipGraph and
    replayed. Reports aggregate Mrays/s of the replay, the per-scene latency of tracing each scene alone, and the
    agreement of every hit id with the FP64 path."""
    import torch

    from blazert_amd import PRECISION_FP32_TRAV_FP64_HIT, PRECISION_FP64, Batch, Scene, workloads as W
    from blazert_amd.autotune import tune_batch, tune_leaf

    specs = W.uneven_batch()
    items, leaves = [], []
    t0 = time.perf_counter()
    for sp in specs:
        s = Scene()
        s.add_mesh(sp["vertices"], sp["faces"])
        O = tuple(torch.from_numpy(np.ascontiguousarray(sp["origins"][:, k])).to(dev) for k in range(3))
        D = tuple(torch.from_numpy(np.ascontiguousarray(sp["directions"][:, k])).to(dev) for k in range(3))
        m = min(len(sp["origins"]), 1 << 15)
        r = tune_leaf(s, tuple(c[:m] for c in O), tuple(c[:m] for c in D), precision=PRECISION_FP32_TRAV_FP64_HIT,
                      leaf_sizes=(2, 4, 8), repeats=2)
        leaves.append(r["min_leaf_primitives"])
        items.append((s, O, D))
    prep_s = time.perf_counter() - t0
    n_total = sum(len(sp["origins"]) for sp in specs)
    batch = Batch(items, precision=PRECISION_FP32_TRAV_FP64_HIT)
    tuned = tune_batch(batch, repeats=2)
    stream = torch.cuda.Stream()
    batch.capture(stream)
    torch.cuda.synchronize()
    for _ in range(args.warmup):
        batch.replay(stream)
    stream.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ms = []
    for _ in range(args.steps):
        e0.record(stream)
        batch.replay(stream)
        e1.record(stream)
        e1.synchronize()
        ms.append(e0.elapsed_time(e1))
    replay_ms = float(np.mean(ms))
    # each scene alone (same precision, default launch): its latency when not batched
    alone = []
    for s, O, D in items:
        t, _ = time_launch(lambda: s.intersect(O, D, precision=PRECISION_FP32_TRAV_FP64_HIT), 3, 1)
        alone.append(t)
    agree_n, agree_same, max_rel = 0, 0, 0.0
    for (s, O, D), h in zip(items, batch.hits):
        r = s.intersect(O, D, precision=PRECISION_FP64)
        same = (r["prim_id"] == h["prim_id"]) & (r["geom_id"] == h["geom_id"])
        agree_n += same.numel()
        agree_same += int(same.sum().item())
        hit = same & (r["geom_id"] != -1)
        if hit.any():
            max_rel = max(max_rel, float(((r["t"] - h["t"]).abs() / r["t"].abs())[hit].max().item()))
    log(f"batch: {len(specs)} scenes, {n_total} rays, replay {replay_ms:.3f} ms -> {n_total / replay_ms / 1e3:.1f} "
        f"Mrays/s; scenes alone sum {sum(alone):.3f} ms")
    return {"workload": f"{len(specs)} triangle scenes ({sum(len(sp['faces']) for sp in specs)} tris in total, "
                        f"1k-500k each) x {n_total} incoherent rays (16k-2M per scene), split precision, one "
                        "hipGraph replay",
            "mrays_s": round(n_total / replay_ms / 1e3, 3), "replay_ms": round(replay_ms, 4),
            "scenes_alone_sum_ms": round(float(sum(alone)), 4),
            "per_scene_alone_ms": {"min": round(float(min(alone)), 4), "median": round(float(np.median(alone)), 4),
                                   "max": round(float(max(alone)), 4)},
            "tuned_launch": {k: tuned[k] for k in ("block_size", "refill", "stack_ring")},
            "tuned_vs_default_ms": [round(tuned["best_ms"], 4), round(tuned["default_ms"], 4)],
            "tuned_leaf_sizes": leaves, "prep_s": round(prep_s, 2),
            "hit_id_agreement_vs_fp64": agree_same / max(agree_n, 1), "max_rel_dt_agreeing": max_rel}


def adjoint_leg(args, dev) -> dict:
    """Differentiable 3-bounce mirror tracing: the 1,310,720-triangle icosphere inside a 12-triangle box (one merged
    mesh), 4,194,304 rays, loss = sum over bounces of w * t with w ~ U(0,1). Times the forward (3 traced bounces) and
    the FP64 adjoint (dL/dvertex, dL/dorigin, dL/ddirection), and checks the adjoint against the oracle's analytic
    gradient on a ray sample and against central differences on 64 vertices (face ids fixed, h = 3e-8) and on
    64 rays (re-traced, h = 1e-6), relative to each gradient entry."""
    import torch

    from blazert_amd import Scene, workloads as W
    from blazert_amd.mirror import mirror_adjoint, mirror_trace
    from oracle import oracle as OR

    B = 3
    V, F = W.mirror_box_scene(8)
    org, dirs = W.mirror_rays()
    n = len(org)
    w = W.mirror_weights(n, B)
    scene = Scene()
    scene.add_mesh(V, F)
    t0 = time.perf_counter()
    scene.commit()
    scene.upload()
    build_s = time.perf_counter() - t0

    def cols(a):
        return tuple(torch.from_numpy(np.ascontiguousarray(a[:, k])).to(dev) for k in range(3))

    O, D = cols(org), cols(dirs)
    Wt = torch.from_numpy(w).to(dev)
    fwd_ms, _ = time_launch(lambda: mirror_trace(scene, O, D, B), args.steps, args.warmup)
    fwd32_ms, _ = time_launch(lambda: mirror_trace(scene, O, D, B, precision=1), args.steps, args.warmup)
    out32 = mirror_trace(scene, O, D, B, precision=1)
    out = mirror_trace(scene, O, D, B)
    torch.cuda.synchronize()
    fwd32_same = float((out32["prim_id"] == out["prim_id"]).all(dim=0).double().mean().item())
    del out32
    prim = out["prim_id"]
    gv = torch.zeros((len(V), 3), dtype=torch.float64, device=dev)

    def step():
        gv.zero_()
        return mirror_adjoint(scene, O, D, prim, Wt, None, grad_vertices=gv)

    adj_ms, _ = time_launch(step, args.steps, args.warmup)
    _, go, gd = step()
    torch.cuda.synchronize()
    log(f"mirror: forward {fwd_ms:.3f} ms ({B} bounces; split precision {fwd32_ms:.3f} ms, same face ids on "
        f"{fwd32_same:.6f} of the rays), adjoint {adj_ms:.3f} ms -> {n / adj_ms / 1e3:.1f} Mrays/s")
    pid = prim.cpu().numpy().view(np.uint32)
    gv_h, go_h, gd_h = gv.cpu().numpy(), go.cpu().numpy(), gd.cpu().numpy()
    threads = len(os.sched_getaffinity(0))

    # forward parity on a sample (bit-exact against the oracle's mirror tracing)
    osc = OR.Scene()
    osc.add_mesh(V, F)
    osc.commit()
    ns = 1 << 14
    ref = OR.mirror_trace(osc, V, F, org[:ns], dirs[:ns], B, 1e-9, nthreads=threads)
    tt = out["t"][:, :ns].cpu().numpy()
    pos = np.moveaxis(out["position"][:, :, :ns].cpu().numpy(), 0, -1)
    fwd_exact = bool(np.array_equal(tt.view(np.uint64), ref["t"].view(np.uint64))
                     and np.array_equal(pid[:, :ns], ref["prim_id"])
                     and np.array_equal(pos.view(np.uint64), ref["position"].view(np.uint64)))

    # analytic CPU gradient on a ray sample: GPU adjoint of the same sample against the oracle
    na = 1 << 18
    sub = slice(0, na)
    gvs, gos, gds = mirror_adjoint(scene, tuple(c[sub] for c in O), tuple(c[sub] for c in D),
                                   prim[:, sub].contiguous(), Wt[:, sub].contiguous(), None)
    torch.cuda.synchronize()
    rv, ro, rd = OR.mirror_grad(V, F, org[sub], dirs[sub], pid[:, sub], w[:, sub], nthreads=threads)
    an_err = max(float(np.abs(a.cpu().numpy() - b).max() / np.abs(b).max())
                 for a, b in ((gvs, rv), (gos, ro), (gds, rd)))

    # central differences on 64 vertices with the face ids fixed (only rays touching the vertex change), relative
    # to each entry. The step is h = 3e-8: at 1e-6 the truncation error (h^2 times the third derivative) reaches
    # 1e-4 relative on vertices whose gradient nearly cancels, and below 1e-8 the rounding of the losses dominates.
    h_v = 3e-8
    rng = np.random.default_rng(5)
    valid = pid != OR.INVALID
    touched = np.unique(F[pid[valid]].reshape(-1))
    verts = np.concatenate([rng.choice(touched[touched < len(V) - 8], 60, replace=False), len(V) - 1 - np.arange(4)])
    fd_err, fd_big_err, fd_over = 0.0, 0.0, 0

    def central(vi, k, hh, idx):
        Vp, Vm = V.copy(), V.copy()
        Vp[vi, k] += hh
        Vm[vi, k] -= hh
        lp = OR.mirror_loss(Vp, F, org[idx], dirs[idx], pid[:, idx], w[:, idx], nthreads=threads)
        lm = OR.mirror_loss(Vm, F, org[idx], dirs[idx], pid[:, idx], w[:, idx], nthreads=threads)
        return (lp - lm).sum() / (2 * hh)

    for vi in verts:
        k = int(rng.integers(3))
        faces_v = np.nonzero((F == vi).any(axis=1))[0]
        fmask = np.zeros(len(F) + 1, bool)
        fmask[faces_v] = True
        idx = np.nonzero(fmask[np.where(valid, pid, len(F))].any(axis=0))[0]
        g = gv_h[vi, k]
        den = abs(g) if g != 0 else 1.0
        e = abs(central(vi, k, h_v, idx) - g) / den
        fd_err = max(fd_err, e)
        fd_over += int(e > 1e-6)
        fd_big_err = max(fd_big_err, abs(central(vi, k, 1e-6, idx) - g) / den)

    # central differences on 64 rays, every perturbed ray re-traced on the GPU
    h_r = 1e-6
    ridx = np.arange(0, n, n // 64)[:64]
    pert_o, pert_d = [], []
    for which in (0, 1):
        for k in range(3):
            for sgn in (1.0, -1.0):
                a, b = org[ridx].copy(), dirs[ridx].copy()
                (a if which == 0 else b)[:, k] += sgn * h_r
                pert_o.append(a)
                pert_d.append(b)
    po = torch.from_numpy(np.concatenate(pert_o)).to(dev)
    pd = torch.from_numpy(np.concatenate(pert_d)).to(dev)
    pr = mirror_trace(scene, po, pd, B)
    torch.cuda.synchronize()
    pp = pr["prim_id"].cpu().numpy().view(np.uint32).reshape(B, 12, 64)
    pt = pr["t"].cpu().numpy().reshape(B, 12, 64)
    L = np.where(pp != OR.INVALID, w[:, ridx][:, None] * pt, 0).sum(0)
    ray_err, flipped = 0.0, 0
    for j in range(6):
        which, k = divmod(j, 3)
        same = (pp[:, 2 * j] == pid[:, ridx]).all(0) & (pp[:, 2 * j + 1] == pid[:, ridx]).all(0)
        flipped += int((~same).sum())
        fd = (L[2 * j] - L[2 * j + 1]) / (2 * h_r)
        g = (go_h if which == 0 else gd_h)[ridx, k]
        for r in np.nonzero(same)[0]:
            ray_err = max(ray_err, abs(fd[r] - g[r]) / (abs(g[r]) if g[r] != 0 else 1.0))

    hits = int(valid.sum())
    # algorithmic bytes: per ray 48 B of rays, 12 B of face ids, 24 B of dL/dt and 48 B of ray gradients; per bounce
    # hit a 12 B face record and 72 B of vertices, read twice (forward recompute and reverse sweep)
    adj_bytes = 132 * n + 168 * hits
    adj_ach = adj_bytes / (adj_ms * 1e-3) / 1e9
    sph_corners = int((valid & (pid < len(F) - 12)).sum()) * 3
    log(f"adjoint: {adj_ach:.0f} GB/s algorithmic ({adj_ach / HBM_PEAK_GBS:.3f} of HBM); {sph_corners} sphere-vertex "
        f"corner gradients scattered with FP64 atomics")
    return {"mrays_s": round(n / adj_ms / 1e3, 3), "unit": "Mrays/s", "ms_per_step": round(adj_ms, 4),
            "roofline": {"bound": "hbm", "achieved": round(adj_ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": round(adj_ach / HBM_PEAK_GBS, 4), "bytes_per_launch": adj_bytes,
                         "bytes": "132 B per ray + 168 B per bounce hit"},
            "sphere_corner_atomics": {"corners": sph_corners, "fp64_adds": 3 * sph_corners,
                                      "note": "corners whose vertex finds no slot in the block's LDS table: three "
                                              "lanes add x, y, z of one vertex in one request"},
            "forward_ms": round(fwd_ms, 4), "forward_bounce_mrays_s": round(B * n / fwd_ms / 1e3, 3),
            "forward_split_precision_ms": round(fwd32_ms, 4),
            "forward_split_precision_same_face_ids": fwd32_same,
            "fwd_plus_adjoint_mrays_s": round(n / (fwd_ms + adj_ms) / 1e3, 3),
            "workload": f"displaced icosphere (1,310,720 tris) + 12-tri box, {n} rays x {B} mirror bounces, "
                        "loss = sum w*t, w ~ U(0,1)",
            "bounce_hits": hits, "box_bounce_hits": int((valid & (pid >= len(F) - 12)).sum()),
            "commit_s": round(build_s, 3),
            "forward_bit_exact_vs_oracle": fwd_exact, "forward_sample_rays": ns,
            "max_rel_err_vs_analytic_cpu": an_err, "analytic_sample_rays": na,
            "fd_vertices_max_rel_err": float(fd_err), "fd_vertices": 64, "fd_vertices_above_1e-6": fd_over,
            "fd_bar": {"bar": 1e-6, "relative_to": "|g| of each entry", "h_vertices": h_v, "h_rays": h_r,
                       "met": bool(fd_over == 0 and ray_err <= 1e-6),
                       "vertices_max_rel_err_at_h_1e-6": float(fd_big_err)},
            "fd_rays_max_rel_err": float(ray_err), "fd_rays": 64, "fd_ray_perturbations_flipped": flipped}


def device_commit_leg(v, f, host_scene, host_s) -> dict:
    """commit() built on the GPU (brt_scene_commit_device) for the same mesh: one warm-up build, then three timed
    ones, each into a fresh scene (the median is reported: a commit's host-side allocations vary from call to call);
    the device arrays must equal the host-built scene's byte for byte."""
    import torch

    from blazert_amd import Scene, capi
    warm = Scene()
    warm.add_mesh(v, f)
    warm.commit(device=True)
    warm.close()
    times, same = [], True
    for _ in range(3):
        d = Scene()
        d.add_mesh(v, f)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        d.commit(device=True)
        times.append(time.perf_counter() - t0)
        same = same and all(np.array_equal(host_scene.read_device_array(0, a), d.read_device_array(0, a))
                            for a in (capi.ARRAY_NODES64, capi.ARRAY_NODES4, capi.ARRAY_PRIMITIVES,
                                      capi.ARRAY_PRIM_IDS))
        same = same and d.build_stats(0)["built_on_device"] == 1
        d.close()
    dev_s = float(np.median(times))
    log(f"device commit {dev_s * 1e3:.1f} ms (median of {len(times)}, min {min(times) * 1e3:.1f}; host "
        f"{host_s * 1e3:.1f} ms), identical arrays: {same}")
    return {"device_ms": round(dev_s * 1e3, 2), "device_ms_min": round(min(times) * 1e3, 2),
            "host_ms": round(host_s * 1e3, 2), "speedup": round(host_s / dev_s, 2), "identical_device_arrays": same,
            "host_threads": len(os.sched_getaffinity(0))}


def distance_leg(args, dev, v, f, scene, osc) -> dict:
    """Distance queries on the 1,310,720-triangle icosphere: the closest surface of 1,048,576 points uniform in
    [-2,2]^3 (brt_closest_surface), and distance_to_surface for 4,194,304 random (point, triangle) pairs. The first
    256 closest-surface results are compared bit for bit with the oracle's brute-force minimum."""
    import torch

    from blazert_amd import workloads as W
    n = 1 << 20
    pts = W.incoherent_rays(n, 61, 2.0)[0]
    P = torch.from_numpy(pts).to(dev)
    ms_c, _ = time_launch(lambda: scene.closest_surface(P), max(3, args.steps // 4), 1)
    out = scene.closest_surface(P)
    torch.cuda.synchronize()
    ns = 256
    ref = osc.closest_surface(pts[:ns], nthreads=len(os.sched_getaffinity(0)))
    exact = bool(np.array_equal(out["distance"][:ns].cpu().numpy().view(np.uint64), ref["distance"].view(np.uint64))
                 and np.array_equal(out["prim_id"][:ns].cpu().numpy().view(np.uint32), ref["prim_id"]))
    npair = 1 << 22
    qp = torch.from_numpy(W.incoherent_rays(npair, 62, 2.0)[0]).to(dev)
    ids = torch.from_numpy((W.Rng(63).uniform(npair) * len(f)).astype(np.int32)).to(dev)
    ms_p, _ = time_launch(lambda: scene.distance_to_surface(0, qp, ids), args.steps, 2)
    log(f"closest surface {n / ms_c / 1e3:.2f} Mpoints/s (exact on {ns}: {exact}), "
        f"pairs {npair / ms_p / 1e3:.1f} Mpairs/s")
    # per pair: the point (24 B) and primitive id (4 B) in, the leaf position of the id (4 B), the triangle's three
    # vertices in leaf order (72 B), the distance out (8 B)
    pair_bytes = 24 + 4 + 4 + 72 + 8
    ach = pair_bytes * npair / (ms_p * 1e-3) / 1e9
    return {"closest_surface_mpoints_s": round(n / ms_c / 1e3, 3), "closest_ms": round(ms_c, 3), "points": n,
            "closest_bit_exact_vs_bruteforce_sample": exact, "sample": ns,
            "pairs_mpairs_s": round(npair / ms_p / 1e3, 2), "pairs": npair,
            "pairs_roofline": {"bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                               "frac": round(ach / HBM_PEAK_GBS, 4), "bytes_per_pair": pair_bytes}}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(osc, org, dirs, ns, threads) -> dict:
    """The oracle (the reference-exact C restatement of intersect1, gcc -O3 for generic x86-64, no FMA contraction)
    timed on this host: plain intersect1 (hit records only, no traversal trace, no normals), OpenMP
    schedule(dynamic, 1024) over rays on every core of this process's affinity, and on one core over the first
    quarter of the sample."""
    t0 = time.perf_counter()
    osc.intersect_batch(org[:ns], dirs[:ns], normals=False, trace=False, nthreads=threads)
    all_s = time.perf_counter() - t0
    n1 = max(1, ns // 4)
    t0 = time.perf_counter()
    osc.intersect_batch(org[:n1], dirs[:n1], normals=False, trace=False, nthreads=1)
    one_s = time.perf_counter() - t0
    log(f"cpu baseline: {ns / all_s / 1e6:.2f} Mrays/s on {threads} threads, {n1 / one_s / 1e6:.3f} on 1")
    return {"value": round(ns / all_s / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "one_core_value": round(n1 / one_s / 1e6, 4), "cpu_model": cpu_model(),
            "sample": f"{ns} of the incoherent rays on {threads} threads ({all_s:.2f} s wall) and the first {n1} "
                      f"on 1 thread ({one_s:.2f} s); oracle intersect1 without trace or normals, gcc -O3 generic "
                      "x86-64, OpenMP schedule(dynamic,1024)"}


def kernel_sources_hash() -> str:
    import hashlib
    h = hashlib.sha256()
    for name in ("kernels.hip", "engine.h", "layout.h"):
        h.update(open(os.path.join(ROOT, "blazert_amd", "csrc", name), "rb").read())
    h.update(open(os.path.join(ROOT, "include", "blazert_amd", "device", "exact.h"), "rb").read())
    return h.hexdigest()[:16]


def measured_traffic():
    """HBM bytes per launch of the FP64 and split traversal kernels from profiles/traffic_fp64_icosphere.json
    (rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of this workload, copy-calibrated). Reported only when the kernel
    sources hash to the value recorded with the measurement; otherwise null, since the figure would describe other
    code."""
    tfile = os.path.join(ROOT, "profiles", "traffic_fp64_icosphere.json")
    if not os.path.exists(tfile):
        return None, None, "no measurement (profiles/traffic_fp64_icosphere.json missing)"
    rec = json.load(open(tfile))
    cur = kernel_sources_hash()
    if rec.get("kernel_sources_sha16") != cur:
        return None, None, (f"profiles/traffic_fp64_icosphere.json was measured on kernel sources "
                            f"{rec.get('kernel_sources_sha16')}, these are {cur}: not reported")
    return rec.get("traffic_bytes"), rec.get("split", {}).get("traffic_bytes"), (
        "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of this kernel on this workload, copy-calibrated: "
        f"profiles/traffic_fp64_icosphere.json (kernel sources {cur})")


def float_scene_leg(args, dev, v, f, org, dirs) -> dict:
    """blazert::Scene<float> on the same workload: the icosphere and the 4,194,304 incoherent rays in FP32
    (brt_intersect_f32); the first 65,536 rays are compared bit for bit with the FP32 oracle."""
    import torch

    from blazert_amd import Scene
    from oracle import oracle as OR
    s = Scene(dtype=np.float32)
    s.add_mesh(v, f)
    t0 = time.perf_counter()
    s.commit()
    build_s = time.perf_counter() - t0
    s.upload()
    o32 = np.ascontiguousarray(org, dtype=np.float32)
    d32 = np.ascontiguousarray(dirs, dtype=np.float32)
    O = tuple(torch.from_numpy(np.ascontiguousarray(o32[:, k])).to(dev) for k in range(3))
    D = tuple(torch.from_numpy(np.ascontiguousarray(d32[:, k])).to(dev) for k in range(3))
    n = len(org)
    ms, _ = time_launch(lambda: s.intersect(O, D), args.steps, args.warmup)
    h = s.intersect(O, D)
    torch.cuda.synchronize()
    ns = 1 << 16
    osc = OR.SceneF32()
    osc.add_mesh(v, f)
    osc.commit()
    ref = osc.intersect_batch(o32[:ns], d32[:ns], normals=False, trace=True, nthreads=len(os.sched_getaffinity(0)))
    ok = ref["trace"]["overflow"] == 0
    exact = bool(np.array_equal(h["t"][:ns].cpu().numpy()[ok].view(np.uint32), ref["t"][ok].view(np.uint32))
                 and np.array_equal(h["prim_id"][:ns].cpu().numpy().view(np.uint32)[ok], ref["prim_id"][ok]))
    s.close()
    log(f"Scene<float> {n / ms / 1e3:.1f} Mrays/s ({ms:.3f} ms), exact on {ns}: {exact}")
    return {"value": round(n / ms / 1e3, 3), "unit": "Mrays/s", "ms_per_step": round(ms, 4), "commit_s": round(build_s, 3),
            "bit_exact_vs_f32_oracle_sample": exact, "sample": ns}


def user_geometry_leg(args, dev) -> dict:
    """A user-defined collection (examples/user_geometry, include/blazert_amd/user_geometry.h): 1,048,576 spheres in
    [-40, 40]^3 and 4,194,304 incoherent rays from inside the cloud, traced by the header's one-ray-per-thread
    traversal with the example's sphere intersector, beside the engine's built-in sphere path on the same tree; the
    two hit records are compared bit for bit over the whole batch."""
    import ctypes as C

    import torch

    from blazert_amd import BVH, Scene, capi
    from blazert_amd import workloads as W
    lib = C.CDLL(os.path.join(ROOT, "examples", "user_geometry", "libuser_geometry.so"))
    P, U, U64 = C.c_void_p, C.c_uint, C.c_uint64
    lib.ug_spheres_intersect.argtypes = [P, P, P, U, C.POINTER(capi.Rays), C.POINTER(capi.Hits), U64, U, P]
    c, r = W.sphere_field(1 << 20, seed=5)
    c = np.ascontiguousarray(c * 4.0)
    s = Scene()
    s.add_spheres(c, r)
    s.commit()
    s.upload()
    t0 = time.perf_counter()
    bvh = BVH(box_min=c - r[:, None], box_max=c + r[:, None], centers=c)
    bvh.device_view()
    build_s = time.perf_counter() - t0
    org, dirs = W.incoherent_rays(1 << 22, 6, 40.0)
    n = len(org)
    O = tuple(torch.from_numpy(np.ascontiguousarray(org[:, k])).to(dev) for k in range(3))
    D = tuple(torch.from_numpy(np.ascontiguousarray(dirs[:, k])).to(dev) for k in range(3))
    cd, rd = torch.from_numpy(c).to(dev), torch.from_numpy(np.ascontiguousarray(r)).to(dev)
    hu = {k: torch.empty(n, dtype=torch.float64, device=dev) for k in ("t", "u", "v")}
    hu["prim_id"] = torch.empty(n, dtype=torch.int32, device=dev)
    hu["geom_id"] = torch.empty(n, dtype=torch.int32, device=dev)
    R = capi.Rays(*[x.data_ptr() for x in O], *[x.data_ptr() for x in D], None, None)
    H = capi.Hits(hu["t"].data_ptr(), hu["u"].data_ptr(), hu["v"].data_ptr(), hu["prim_id"].data_ptr(),
                  hu["geom_id"].data_ptr(), None, None, None)

    def user():
        rc = lib.ug_spheres_intersect(bvh.handle, cd.data_ptr(), rd.data_ptr(), len(r), C.byref(R), C.byref(H), n, 0,
                                      C.c_void_p(torch.cuda.current_stream().cuda_stream))
        if rc != 0:
            raise RuntimeError(f"ug_spheres_intersect: {rc}")

    ms_user, _ = time_launch(user, args.steps, args.warmup)
    ms_builtin, _ = time_launch(lambda: s.intersect(O, D), args.steps, args.warmup)
    hb = s.intersect(O, D)
    user()
    torch.cuda.synchronize()
    exact = bool(torch.equal(hu["t"].view(torch.int64), hb["t"].view(torch.int64))
                 and torch.equal(hu["prim_id"], hb["prim_id"]) and torch.equal(hu["geom_id"], hb["geom_id"]))
    hits = int((hb["prim_id"] != -1).sum())
    s.close()
    bvh.close()
    # algorithmic bytes per ray (both paths read the same tree and sphere records): the reference traversal's node
    # visits x 56 B + sphere tests x 36 B + 96 B of ray and hit, from the oracle's replay on the first 2^18 rays
    from oracle import oracle as OR
    osph = OR.Scene()
    osph.add_spheres(c, r)
    osph.commit()
    nsmp = 1 << 18
    tr = osph.intersect_batch(org[:nsmp], dirs[:nsmp], normals=False, trace=True,
                              nthreads=len(os.sched_getaffinity(0)))["trace"]
    bpr = S_RAY_HIT + float(tr["node_visits"].mean()) * S_NODE64 + float(tr["prim_tests"].mean()) * 36
    roof = {k: {"bound": "hbm", "achieved": round(bpr * n / (t * 1e-3) / 1e9, 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(bpr * n / (t * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
            for k, t in (("user", ms_user), ("builtin", ms_builtin))}
    log(f"user collection {n / ms_user / 1e3:.1f} Mrays/s ({ms_user:.3f} ms) vs built-in spheres "
        f"{n / ms_builtin / 1e3:.1f} Mrays/s, identical: {exact}")
    return {"value": round(n / ms_user / 1e3, 3), "unit": "Mrays/s", "ms_per_step": round(ms_user, 4),
            "builtin_spheres_mrays_s": round(n / ms_builtin / 1e3, 3), "builtin_ms": round(ms_builtin, 4),
            "identical_to_builtin_all_rays": exact, "hits": hits, "tree_build_s": round(build_s, 3),
            "bytes_per_ray": round(bpr, 1), "roofline_user": roof["user"], "roofline_builtin": roof["builtin"],
            "workload": "1,048,576 spheres (r in [0.1, 0.5]) in [-40, 40]^3, 4,194,304 incoherent rays from inside, "
                        "example sphere intersector through include/blazert_amd/user_geometry.h"}


def ceilings(scene_bytes: int, lanes: int, adjoint_rows: int, adjoint_adds: int) -> dict:
    """The memory ceilings the kernels are compared with, measured on this GPU by blazert_amd/libbrt_microbench.so
    (csrc/microbench.hip): `lanes` dependent chains of random 128-byte lines in a buffer of the scene's size (the
    traversal's pattern: each record's address comes from the previous record), independent random 128-byte
    gathers (plain and non-temporal loads), and random FP64 atomic adds into the adjoint's (vertices, 3) gradient,
    one component per lane and as lane triples (x, y, z of one vertex in three adjacent lanes)."""
    import ctypes as C

    from blazert_amd import build as B
    lib = C.CDLL(B.MICROBENCH_LIB)
    U64, U32, DP = C.c_uint64, C.c_uint32, C.POINTER(C.c_double)
    lib.brt_mb_line_chase.argtypes = [U64, U32, U32, DP]
    lib.brt_mb_line_gather.argtypes = [U64, U64, C.c_int, DP]
    lib.brt_mb_fp64_atomics.argtypes = [U64, U64, C.c_int, DP]
    r = {}

    def run(name, fn, *a):
        x = C.c_double(0.0)
        rc = fn(*a, C.byref(x))
        if rc != 0:
            raise RuntimeError(f"microbenchmark {name} failed: {rc}")
        r[name] = round(x.value, 1)

    run("dependent_line_chains_gbs", lib.brt_mb_line_chase, scene_bytes, lanes, 200)
    run("random_line_gather_gbs", lib.brt_mb_line_gather, scene_bytes, 1 << 26, 0)
    run("random_line_gather_nontemporal_gbs", lib.brt_mb_line_gather, scene_bytes, 1 << 26, 1)
    adds = adjoint_adds - adjoint_adds % 3
    run("fp64_atomics_per_lane_gadds", lib.brt_mb_fp64_atomics, adjoint_rows, adds, 0)
    run("fp64_atomics_lane_triples_gadds", lib.brt_mb_fp64_atomics, adjoint_rows, adds, 1)
    r["config"] = {"buffer_bytes": scene_bytes, "chains": lanes, "chain_steps": 200, "gathers": 1 << 26,
                   "atomic_rows": adjoint_rows, "atomic_adds": adds}
    return r


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-rays", type=int, default=1 << 22)
    ap.add_argument("--no-adjoint", action="store_true", help="skip the mirror-tracing adjoint leg")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-primitive split-precision leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the uneven multi-scene batch leg")
    ap.add_argument("--no-distance", action="store_true", help="skip the distance-query leg")
    ap.add_argument("--no-float", action="store_true", help="skip the Scene<float> leg")
    ap.add_argument("--no-user", action="store_true", help="skip the user-defined collection leg")
    ap.add_argument("--microbench", action="store_true",
                    help="only run the ceiling microbenchmarks (blazert_amd/libbrt_microbench.so) and print them")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy (scale factor for FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported in this build "
              "(rays are independent: a multi-GPU run would replicate the scene on every GPU and shard the rays)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    if args.calibrate_copy:
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda:0")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"calibrate_copy_bytes": a.numel() * 8, "copies": 3}))
        return

    # the scene's device footprint (nodes, records, ids: about 190 MB), the FP64 kernel's grid, and the mirror
    # scene's vertices with the bench adjoint's count of sphere-vertex FP64 adds
    CEIL = dict(scene_bytes=192 << 20, lanes=262144, adjoint_rows=655370, adjoint_adds=24221979)
    if args.microbench:
        print(json.dumps(ceilings(**CEIL)))
        return

    from blazert_amd import PRECISION_FP32_TRAV_FP64_HIT, PRECISION_FP64, Scene, workloads as W

    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    v, f = W.icosphere(8)
    org, dirs = W.icosphere_incoherent_rays()
    corg, cdirs = W.icosphere_coherent_rays()
    log(f"inputs {time.perf_counter() - t0:.1f}s: {len(f)} tris, {len(org)} incoherent, {len(corg)} coherent")

    scene = Scene()
    scene.add_mesh(v, f)
    t0 = time.perf_counter()
    scene.commit()
    build_s = time.perf_counter() - t0
    stats = scene.build_stats(0)
    scene.upload()
    log(f"commit {build_s:.2f}s depth {stats['max_tree_depth']} leaves {stats['leaf_nodes']}")

    def cols(a):  # structure-of-arrays columns on the device, made once outside the timed region
        return tuple(torch.from_numpy(np.ascontiguousarray(a[:, k])).to(dev) for k in range(3))

    O, D = cols(org), cols(dirs)
    CO, CD = cols(corg), cols(cdirs)
    n = len(org)

    ms64, wall64 = time_launch(lambda: scene.intersect(O, D, precision=PRECISION_FP64), args.steps, args.warmup)
    log(f"fp64 incoherent {ms64:.3f} ms/step -> {n / ms64 / 1e3:.1f} Mrays/s")
    ms32, _ = time_launch(lambda: scene.intersect(O, D, precision=PRECISION_FP32_TRAV_FP64_HIT), args.steps,
                          args.warmup)
    log(f"fp32-trav incoherent {ms32:.3f} ms/step -> {n / ms32 / 1e3:.1f} Mrays/s")
    cms64, _ = time_launch(lambda: scene.intersect(CO, CD, precision=PRECISION_FP64), args.steps, args.warmup)
    cms32, _ = time_launch(lambda: scene.intersect(CO, CD, precision=PRECISION_FP32_TRAV_FP64_HIT), args.steps,
                           args.warmup)
    nc = len(corg)
    log(f"coherent fp64 {nc / cms64 / 1e3:.1f} / fp32-trav {nc / cms32 / 1e3:.1f} Mrays/s")
    # the same coherent rays with 32-index ray-pool claims (the pool_size launch option the autotuner tries)
    cp64, _ = time_launch(lambda: scene.intersect(CO, CD, precision=PRECISION_FP64, pool_size=32), args.steps,
                          args.warmup)
    cp32, _ = time_launch(lambda: scene.intersect(CO, CD, precision=PRECISION_FP32_TRAV_FP64_HIT, pool_size=32),
                          args.steps, args.warmup)
    log(f"coherent, 32-ray claims: fp64 {nc / cp64 / 1e3:.1f} / fp32-trav {nc / cp32 / 1e3:.1f} Mrays/s")

    # agreement of the split-precision path with the FP64 path on the full incoherent set
    h64 = scene.intersect(O, D, precision=PRECISION_FP64)
    h32 = scene.intersect(O, D, precision=PRECISION_FP32_TRAV_FP64_HIT)
    torch.cuda.synchronize()
    agree, max_rel, dis_idx, dis_other = split_agreement(h64, h32)

    out = {
        "metric": METRIC,
        "value": round(n / ms64 / 1e3, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms64, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "displaced icosphere (8 subdivisions, 1,310,720 tris) x 4,194,304 incoherent rays "
                               "(origins U[-2,2]^3, directions uniform on S^2), FP64 intersect1",
                   "triangles": int(len(f)), "rays": n, "commit_s": round(build_s, 3),
                   "tree_depth": stats["max_tree_depth"]},
        "split_precision": {"value": round(n / ms32 / 1e3, 3), "unit": "Mrays/s", "ms_per_step": round(ms32, 4),
                            "speedup_vs_fp64": round(ms64 / ms32, 3), "hit_id_agreement": agree,
                            "max_rel_dt_agreeing": max_rel},
        "coherent_1920x1080": {"fp64_mrays_s": round(nc / cms64 / 1e3, 3),
                               "fp32_trav_mrays_s": round(nc / cms32 / 1e3, 3),
                               "pool_size_32": {"fp64_mrays_s": round(nc / cp64 / 1e3, 3),
                                                "fp32_trav_mrays_s": round(nc / cp32 / 1e3, 3)}},
    }

    # algorithmic bytes per ray from the reference traversal replay (oracle trace on the rays; not timed)
    from oracle import oracle as OR
    osc = OR.Scene()
    osc.add_mesh(v, f)
    osc.commit()
    ns = min(args.cpu_rays, n)
    threads = len(os.sched_getaffinity(0))
    ref = osc.intersect_batch(org[:ns], dirs[:ns], normals=True, trace=True, nthreads=threads)
    tr = ref["trace"]
    visits = float(tr["node_visits"].mean())
    ptests = float(tr["prim_tests"].mean())
    bpr64 = S_RAY_HIT + visits * S_NODE64 + ptests * S_TRI
    g_sub_t = h64["t"][:ns].cpu().numpy()
    g_sub_p = h64["prim_id"][:ns].cpu().numpy().view(np.uint32)
    ok = tr["overflow"] == 0
    exact = bool(np.array_equal(g_sub_t[ok].view(np.uint64), ref["t"][ok].view(np.uint64))
                 and np.array_equal(g_sub_p[ok], ref["prim_id"][ok]))
    achieved = bpr64 * n / (ms64 * 1e-3) / 1e9
    traffic, traffic32, traffic_note = measured_traffic()
    out["roofline"] = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "traffic_source": traffic_note,
                       "bytes_per_ray": round(bpr64, 1), "node_visits_per_ray": round(visits, 2),
                       "tri_tests_per_ray": round(ptests, 2)}
    # split precision: the same work numerator (the reference traversal's node visits and triangle tests) with the
    # split path's node record size, 32 B of FP32 box + topology per node
    bpr32 = S_RAY_HIT + visits * S_NODE32 + ptests * S_TRI
    ach32 = bpr32 * n / (ms32 * 1e-3) / 1e9
    out["split_precision"]["roofline"] = {"bound": "hbm", "achieved": round(ach32, 1), "peak": HBM_PEAK_GBS,
                                          "unit": "GB/s", "frac": round(ach32 / HBM_PEAK_GBS, 4),
                                          "bytes_per_ray": round(bpr32, 1), "traffic": traffic32,
                                          "traffic_source": traffic_note}
    ceil = ceilings(**CEIL)
    out["ceilings"] = ceil
    chase = ceil["dependent_line_chains_gbs"]
    for rf, tb, ms in ((out["roofline"], traffic, ms64), (out["split_precision"]["roofline"], traffic32, ms32)):
        rf["dependent_line_ceiling_gbs"] = chase
        if tb:
            rf["traffic_gbs"] = round(tb / (ms * 1e-3) / 1e9, 1)
            rf["traffic_frac_of_dependent_line_ceiling"] = round(tb / (ms * 1e-3) / 1e9 / chase, 4)
    out["parity"] = {"sample_rays": ns, "fp64_bit_exact_vs_oracle": exact,
                     "reference_stack_overflow_rays": int((~ok).sum())}
    out["split_precision"]["disagreements"] = classify_disagreements(osc, org, dirs, dis_idx, dis_other)
    del h64, h32, O, D, CO, CD
    out["device_commit"] = device_commit_leg(v, f, scene, build_s)
    if not args.no_float:
        out["scene_float"] = float_scene_leg(args, dev, v, f, org, dirs)
    if not args.no_distance:
        out["distance_queries"] = distance_leg(args, dev, v, f, scene, osc)
    if not args.no_user:
        out["user_geometry"] = user_geometry_leg(args, dev)
    if not args.no_mixed:
        out["mixed_split_precision"] = mixed_leg(args, dev)
    if not args.no_batch:
        out["batch_uneven"] = batch_leg(args, dev)
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(args, dev)
        # the sphere-vertex scatter against the measured random FP64 atomic rates
        sc = out["adjoint"]["sphere_corner_atomics"]
        sc["per_lane_ceiling_gadds"] = ceil["fp64_atomics_per_lane_gadds"]
        sc["lane_triples_ceiling_gadds"] = ceil["fp64_atomics_lane_triples_gadds"]
        sc["lane_triples_floor_ms"] = round(sc["fp64_adds"] / ceil["fp64_atomics_lane_triples_gadds"] / 1e6, 4)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(osc, org, dirs, ns, threads)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
