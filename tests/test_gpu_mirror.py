"""GPU mirror tracing and its adjoint (brt_mirror_trace / brt_mirror_adjoint through the C ABI).

Forward: every bounce bit-exact against the CPU oracle (t, u, v, face id, hit position, mirror direction; NaN
where the ray has missed). Adjoint: against the oracle's analytic FP64 gradient (explicit Jacobians, written
independently of the kernel) to 1e-10 of the largest gradient entry, and against central finite differences of
the re-traced loss with every perturbed scene rebuilt, to 1e-6 relative to each gradient entry: vertices with
h = 3e-8 (at 1e-6 the h^2 truncation error of central differences exceeds 1e-6 on a vertex of this set, below 1e-8
the rounding of the losses does), rays with h = 1e-6.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from blazert_amd import workloads as W  # noqa: E402

pytestmark = pytest.mark.gpu

INVALID = 0xFFFFFFFF
H = 1e-6  # ray origins and directions
H_VERTEX = 3e-8


@pytest.fixture(scope="module")
def dev():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    import blazert_amd  # noqa: F401
    return torch.device("cuda:0")


def _gpu_scene(V, F):
    from blazert_amd import Scene
    s = Scene()
    s.add_mesh(V, F)
    s.commit()
    return s


def _oracle_scene(oracle, V, F):
    s = oracle.Scene()
    s.add_mesh(V, F)
    s.commit()
    return s


def _trace(scene, o, d, dev, bounces=3, **kw):
    from blazert_amd.mirror import mirror_trace
    O = torch.from_numpy(np.ascontiguousarray(o)).to(dev)
    D = torch.from_numpy(np.ascontiguousarray(d)).to(dev)
    out = mirror_trace(scene, O, D, bounces, uv=True, **kw)
    torch.cuda.synchronize()
    res = {k: out[k].cpu().numpy() for k in ("t", "u", "v")}
    res["prim_id"] = out["prim_id"].cpu().numpy().view(np.uint32)
    res["position"] = np.moveaxis(out["position"].cpu().numpy(), 0, -1)
    res["direction"] = np.moveaxis(out["direction"].cpu().numpy(), 0, -1)
    return res


def _same_bits(a, b):
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return np.array_equal(a.view(np.uint64), b.view(np.uint64)) if a.dtype == np.float64 else np.array_equal(a, b)


@pytest.mark.parametrize("level,box", [(4, True), (3, False)])
def test_mirror_forward_bit_exact(dev, oracle_mod, level, box):
    print(f"mirror forward level={level} box={box}", flush=True)
    V, F = W.mirror_box_scene(level=level) if box else W.icosphere(level)
    o, d = W.mirror_rays(1 << 15, seed=11)
    g = _trace(_gpu_scene(V, F), o, d, dev)
    r = oracle_mod.mirror_trace(_oracle_scene(oracle_mod, V, F), V, F, o, d, 3, 1e-9)
    for k in ("t", "u", "v", "prim_id", "position", "direction"):
        assert _same_bits(g[k], r[k]), k
    hit = r["prim_id"] != INVALID
    assert hit.all() if box else (~hit).any()


def test_mirror_launch_options_agree(dev, oracle_mod):
    V, F = W.mirror_box_scene(level=3)
    o, d = W.mirror_rays(1 << 14, seed=12)
    s = _gpu_scene(V, F)
    base = _trace(s, o, d, dev)
    for kw in ({"schedule": 1}, {"refill": 1, "block_size": 64}, {"refill": 64, "block_size": 192}):
        g = _trace(s, o, d, dev, **kw)
        for k in ("t", "prim_id", "position", "direction"):
            assert _same_bits(g[k], base[k]), (kw, k)


@pytest.fixture(scope="module")
def adjoint_case(dev, oracle_mod):
    from blazert_amd.mirror import mirror_adjoint, mirror_trace
    V, F = W.mirror_box_scene(level=4)
    n = 1 << 16
    o, d = W.mirror_rays(n, seed=13)
    w = W.mirror_weights(n, 3, seed=14)
    gp = W.Rng(15).uniform(3 * 3 * n).reshape(3, n, 3)
    s = _gpu_scene(V, F)
    O = torch.from_numpy(o).to(dev)
    D = torch.from_numpy(d).to(dev)
    out = mirror_trace(s, O, D, 3)
    Wt = torch.from_numpy(w).to(dev)
    GP = torch.from_numpy(np.ascontiguousarray(np.moveaxis(gp, -1, 0))).to(dev)
    gv, go, gd = mirror_adjoint(s, O, D, out["prim_id"], Wt, GP)
    torch.cuda.synchronize()
    pid = out["prim_id"].cpu().numpy().view(np.uint32)
    return dict(V=V, F=F, o=o, d=d, w=w, gp=gp, pid=pid, gv=gv.cpu().numpy(), go=go.cpu().numpy(),
                gd=gd.cpu().numpy(), scene=s)


def test_adjoint_matches_analytic_cpu_gradient(adjoint_case, oracle_mod):
    c = adjoint_case
    print("adjoint vs oracle gradient, 65536 rays", flush=True)
    gv, go, gd = oracle_mod.mirror_grad(c["V"], c["F"], c["o"], c["d"], c["pid"], c["w"], c["gp"])
    for name, a, b in (("vertices", c["gv"], gv), ("origins", c["go"], go), ("directions", c["gd"], gd)):
        scale = np.abs(b).max()
        assert scale > 0
        err = np.abs(a - b).max() / scale
        assert err <= 1e-10, (name, err)
    # vertices no ray touched get exactly zero
    touched = np.zeros(len(c["V"]), bool)
    touched[c["F"][c["pid"][c["pid"] != INVALID]].reshape(-1)] = True
    assert (c["gv"][~touched] == 0).all()


def test_adjoint_vertex_gradient_vs_retraced_fd(dev, adjoint_case):
    """Central differences of the traced loss with the scene rebuilt per perturbation (64 vertices, one random
    coordinate each, rays whose face sequence flips left out of both sides)."""
    from blazert_amd.mirror import mirror_adjoint, mirror_trace
    c = adjoint_case
    V, F, o, d, w, gp, pid = c["V"], c["F"], c["o"], c["d"], c["w"], c["gp"], c["pid"]
    O = torch.from_numpy(o).to(dev)
    D = torch.from_numpy(d).to(dev)

    def traced(Vx):
        out = mirror_trace(_gpu_scene(Vx, F), O, D, 3)
        torch.cuda.synchronize()
        p = out["prim_id"].cpu().numpy().view(np.uint32)
        ok = p != INVALID
        t = out["t"].cpu().numpy()
        x = np.moveaxis(out["position"].cpu().numpy(), 0, -1)
        return np.where(ok, w * t, 0).sum(0) + np.where(ok[..., None], gp * x, 0).sum((0, 2)), p

    rng = np.random.default_rng(16)
    touched = np.unique(F[pid[pid != INVALID]].reshape(-1))
    verts = np.concatenate([rng.choice(touched, 60, replace=False), len(V) - 1 - np.arange(4)])
    bad = []
    flipped = 0
    for vi in verts:
        k = int(rng.integers(3))
        Vp, Vm = V.copy(), V.copy()
        Vp[vi, k] += H_VERTEX
        Vm[vi, k] -= H_VERTEX
        lp, pp = traced(Vp)
        lm, pm = traced(Vm)
        same = ((pp == pid) & (pm == pid)).all(axis=0)
        flipped += int((~same).sum())
        fd = (lp - lm)[same].sum() / (2 * H_VERTEX)
        if same.all():
            g = c["gv"][vi, k]
        else:
            m = torch.from_numpy(same).to(dev)
            gsub, _, _ = mirror_adjoint(c["scene"], O[m], D[m], torch.from_numpy(pid[:, same].view(np.int32)).to(dev),
                                        torch.from_numpy(np.ascontiguousarray(w[:, same])).to(dev),
                                        torch.from_numpy(np.ascontiguousarray(np.moveaxis(gp[:, same], -1, 0))).to(dev),
                                        want_rays=False)
            g = gsub[vi, k].item()
        if abs(fd - g) > 1e-6 * abs(g):
            bad.append((int(vi), k, fd, g))
    print(f"fd: 64 vertices, {flipped} flipped ray-perturbations excluded", flush=True)
    assert not bad, bad


def test_adjoint_ray_gradient_vs_retraced_fd(dev, adjoint_case):
    from blazert_amd.mirror import mirror_trace
    c = adjoint_case
    s, w, gp = c["scene"], c["w"], c["gp"]
    idx = np.arange(0, len(c["o"]), 1021)[:64]
    o, d = c["o"][idx], c["d"][idx]
    rows = []
    for which in (0, 1):
        for k in range(3):
            for sgn in (1.0, -1.0):
                a, b = o.copy(), d.copy()
                (a if which == 0 else b)[:, k] += sgn * H
                rows.append((a, b))
    A = np.concatenate([r[0] for r in rows])
    Bd = np.concatenate([r[1] for r in rows])
    out = mirror_trace(s, torch.from_numpy(A).to(dev), torch.from_numpy(Bd).to(dev), 3)
    torch.cuda.synchronize()
    p = out["prim_id"].cpu().numpy().view(np.uint32).reshape(3, 12, len(idx))
    t = out["t"].cpu().numpy().reshape(3, 12, len(idx))
    x = np.moveaxis(out["position"].cpu().numpy(), 0, -1).reshape(3, 12, len(idx), 3)
    ww, gg = w[:, idx], gp[:, idx]
    ok = p != INVALID
    L = np.where(ok, ww[:, None] * t, 0).sum(0) + np.where(ok[..., None], gg[:, None] * x, 0).sum((0, 3))
    checked = 0
    for j in range(12 // 2):
        which, k = divmod(j, 3)
        same = (p[:, 2 * j] == c["pid"][:, idx]).all(0) & (p[:, 2 * j + 1] == c["pid"][:, idx]).all(0)
        fd = (L[2 * j] - L[2 * j + 1]) / (2 * H)
        g = (c["go"] if which == 0 else c["gd"])[idx, k]
        for r in np.nonzero(same)[0]:
            assert abs(fd[r] - g[r]) <= 1e-6 * abs(g[r]), (which, k, r, fd[r], g[r])
            checked += 1
    assert checked >= 300


def test_autograd_function(dev, adjoint_case):
    from blazert_amd.mirror import mirror_trace_autograd
    c = adjoint_case
    Vt = torch.from_numpy(c["V"]).to(dev).requires_grad_(True)
    O = torch.from_numpy(c["o"]).to(dev).requires_grad_(True)
    D = torch.from_numpy(c["d"]).to(dev).requires_grad_(True)
    t, pos, pid = mirror_trace_autograd(c["scene"], Vt, O, D, 3)
    ok = pid != -1
    Wt = torch.from_numpy(c["w"]).to(dev)
    GP = torch.from_numpy(c["gp"]).to(dev)
    loss = torch.where(ok, Wt * t, torch.zeros_like(t)).sum() + \
        torch.where(ok[..., None], GP * pos, torch.zeros_like(pos)).sum()
    loss.backward()
    torch.cuda.synchronize()
    scale = np.abs(c["gv"]).max()
    assert np.abs(Vt.grad.cpu().numpy() - c["gv"]).max() <= 1e-10 * scale
    assert np.abs(O.grad.cpu().numpy() - c["go"]).max() <= 1e-10 * np.abs(c["go"]).max()
    assert np.abs(D.grad.cpu().numpy() - c["gd"]).max() <= 1e-10 * np.abs(c["gd"]).max()


def test_split_precision_forward(dev, adjoint_case):
    """Bounces traced with FP32 traversal and FP64 hits give the FP64 forward's face ids, distances and positions on
    every ray whose bounces have no near-tie (a box face is two coplanar triangles: a ray through their shared
    diagonal may take the other triangle, which reflects the same way)."""
    from blazert_amd.mirror import mirror_trace
    c = adjoint_case
    O, D = torch.from_numpy(c["o"]).to(dev), torch.from_numpy(c["d"]).to(dev)
    a = mirror_trace(c["scene"], O, D, 3)
    b = mirror_trace(c["scene"], O, D, 3, precision=1)
    torch.cuda.synchronize()
    pa = a["prim_id"].cpu().numpy().view(np.uint32)
    pb = b["prim_id"].cpu().numpy().view(np.uint32)
    same = (pa == pb).all(axis=0)
    assert same.mean() > 0.999, same.mean()
    ta, tb = a["t"].cpu().numpy()[:, same], b["t"].cpu().numpy()[:, same]
    assert np.array_equal(ta.view(np.uint64), tb.view(np.uint64))
    xa = a["position"].cpu().numpy()[:, :, same]
    xb = b["position"].cpu().numpy()[:, :, same]
    assert np.array_equal(np.nan_to_num(xa).view(np.uint64), np.nan_to_num(xb).view(np.uint64))
