"""Pins the CPU oracle of mirror tracing and of its analytic gradient (oracle/mirror_oracle.c).

The reference has no differentiable path, so nothing of its own test data covers this; the pins are closed forms
(axis-aligned bounces in a box, the mirror law, the loss at the traced state) and central finite differences of the
loss, both with the face ids held fixed and with every perturbed scene rebuilt and re-traced, to 1e-6 relative to
each gradient entry: vertices with h = 3e-8 (at 1e-6 the h^2 truncation error can exceed 1e-6 where a vertex's
contributions nearly cancel), ray origins and directions with h = 1e-6.
"""
import numpy as np
import pytest

from blazert_amd import workloads as W

H = 1e-6  # ray origins and directions
H_VERTEX = 3e-8


def _scene(oracle, V, F):
    s = oracle.Scene()
    s.add_mesh(V, F)
    s.commit()
    return s


def _fd_close(fd, g):
    return abs(fd - g) <= 1e-6 * abs(g)


def test_box_bounces_closed_form(oracle_mod):
    O = oracle_mod
    V, F = W.box_mesh(3.0)
    s = _scene(O, V, F)
    o = np.array([[0.0, 0.0, 0.0], [1.0, 0.5, -2.0], [0.0, 1.0, 0.0]])
    d = np.array([[2.0, 0.0, 0.0], [0.0, 0.0, 1.0], [1.0, 1.0, 0.0]])
    out = O.mirror_trace(s, V, F, o, d, 3, 1e-9)
    # head-on: 3, then across the box and back
    np.testing.assert_array_equal(out["t"][:, 0], [3.0, 6.0, 6.0])
    np.testing.assert_array_equal(out["position"][:, 0], [[3, 0, 0], [-3, 0, 0], [3, 0, 0]])
    np.testing.assert_array_equal(out["direction"][:, 0], [[-1, 0, 0], [1, 0, 0], [-1, 0, 0]])
    np.testing.assert_array_equal(out["t"][:, 1], [5.0, 6.0, 6.0])
    # 45 degrees in the z = 0 plane: y = 3 wall at (2, 3), x = 3 wall at (3, 2), y = -3 wall at (-2, -3)
    np.testing.assert_allclose(out["t"][:, 2], np.sqrt(2) * np.array([2.0, 1.0, 5.0]), rtol=1e-15)
    np.testing.assert_allclose(out["position"][:, 2], [[2, 3, 0], [3, 2, 0], [-2, -3, 0]], atol=1e-14)
    r = np.sqrt(0.5)
    np.testing.assert_allclose(out["direction"][:, 2], [[r, -r, 0], [-r, -r, 0], [-r, r, 0]], atol=1e-15)


def test_mirror_law_and_misses(oracle_mod):
    O = oracle_mod
    V, F = W.icosphere(2)
    s = _scene(O, V, F)
    o, d = W.mirror_rays(3000, seed=4)
    out = O.mirror_trace(s, V, F, o, d, 3, 1e-9)
    pid = out["prim_id"]
    hit = pid != O.INVALID
    assert hit[0].any() and (~hit[0]).any()
    # a miss ends the ray: later bounces miss, positions and directions are NaN
    for b in range(1, 3):
        assert not (hit[b] & ~hit[b - 1]).any()
    assert np.isnan(out["position"][~hit]).all() and np.isnan(out["direction"][~hit]).all()
    assert (out["t"][~hit] == np.finfo(np.float64).max).all()
    # mirror law on every hit: |r| = 1, r.n = -dh.n, r - dh parallel to n
    b, i = np.nonzero(hit)
    f = F[pid[b, i]]
    n = np.cross(V[f[:, 1]] - V[f[:, 0]], V[f[:, 2]] - V[f[:, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    din = np.where(b[:, None] == 0, d[i], 0.0)
    for bb in (1, 2):
        m = b == bb
        din[m] = out["direction"][bb - 1, i[m]]
    dh = din / np.linalg.norm(din, axis=1, keepdims=True)
    r = out["direction"][b, i]
    np.testing.assert_allclose(np.linalg.norm(r, axis=1), 1.0, atol=1e-14)
    np.testing.assert_allclose(np.einsum("ij,ij->i", r, n), -np.einsum("ij,ij->i", dh, n), atol=1e-14)
    np.testing.assert_allclose(np.cross(r - dh, n), 0.0, atol=1e-14)


@pytest.fixture(scope="module")
def small_case(oracle_mod):
    O = oracle_mod
    V, F = W.mirror_box_scene(level=2)
    s = _scene(O, V, F)
    n = 2000
    o, d = W.mirror_rays(n, seed=5)
    out = O.mirror_trace(s, V, F, o, d, 3, 1e-9)
    w = W.mirror_weights(n, 3)
    gp = W.Rng(31).uniform(3 * n * 3).reshape(3, n, 3)
    return V, F, o, d, out, w, gp


def test_loss_at_traced_state(oracle_mod, small_case):
    O = oracle_mod
    V, F, o, d, out, w, gp = small_case
    pid = out["prim_id"]
    ok = pid != O.INVALID
    L = O.mirror_loss(V, F, o, d, pid, w, gp)
    ref = np.where(ok, w * out["t"], 0).sum(0) + np.where(ok[..., None], gp * out["position"], 0).sum((0, 2))
    np.testing.assert_allclose(L, ref, rtol=1e-12)


def test_vertex_gradient_fixed_ids_fd(oracle_mod, small_case):
    O = oracle_mod
    V, F, o, d, out, w, gp = small_case
    pid = out["prim_id"]
    gv, _, _ = O.mirror_grad(V, F, o, d, pid, w, gp)
    touched = np.unique(F[pid[pid != O.INVALID]].reshape(-1))
    vs = np.random.default_rng(1).choice(touched, 16, replace=False)
    vs = np.concatenate([vs, [len(V) - 1]])  # a box corner: thousands of contributions
    for vi in vs:
        for c in range(3):
            Vp, Vm = V.copy(), V.copy()
            Vp[vi, c] += H_VERTEX
            Vm[vi, c] -= H_VERTEX
            fd = ((O.mirror_loss(Vp, F, o, d, pid, w, gp) - O.mirror_loss(Vm, F, o, d, pid, w, gp)).sum()
                  / (2 * H_VERTEX))
            assert _fd_close(fd, gv[vi, c]), (vi, c, fd, gv[vi, c])


def test_ray_gradient_fixed_ids_fd(oracle_mod, small_case):
    O = oracle_mod
    V, F, o, d, out, w, gp = small_case
    pid = out["prim_id"]
    _, go, gd = O.mirror_grad(V, F, o, d, pid, w, gp)
    for i in range(0, 2000, 97):
        for c in range(3):
            for which, g in ((0, go), (1, gd)):
                a, b = (o if which == 0 else d).copy(), (o if which == 0 else d).copy()
                a[i, c] += H
                b[i, c] -= H
                lp = O.mirror_loss(V, F, *((a, d) if which == 0 else (o, a)), pid, w, gp)[i]
                lm = O.mirror_loss(V, F, *((b, d) if which == 0 else (o, b)), pid, w, gp)[i]
                assert _fd_close((lp - lm) / (2 * H), g[i, c]), (i, c, which)


def test_vertex_gradient_retraced_fd(oracle_mod):
    """FD of the traced loss (scene rebuilt and every ray re-traced per perturbation) equals the analytic gradient
    wherever no ray changes its face sequence."""
    O = oracle_mod
    V, F = W.mirror_box_scene(level=1)
    n = 600
    o, d = W.mirror_rays(n, seed=8)
    w = W.mirror_weights(n, 3, seed=9)

    def traced(Vx):
        out = O.mirror_trace(_scene(O, Vx, F), Vx, F, o, d, 3, 1e-9)
        ok = out["prim_id"] != O.INVALID
        return np.where(ok, w * out["t"], 0).sum(0), out["prim_id"]

    _, pid = traced(V)
    gv, _, _ = O.mirror_grad(V, F, o, d, pid, w)
    checked = 0
    for vi in (0, 5, 17, len(V) - 8, len(V) - 1):
        for c in range(3):
            Vp, Vm = V.copy(), V.copy()
            Vp[vi, c] += H_VERTEX
            Vm[vi, c] -= H_VERTEX
            lp, pp = traced(Vp)
            lm, pm = traced(Vm)
            same = ((pp == pid) & (pm == pid)).all(axis=0)
            # rays whose face sequence flipped under the perturbation are left out of both sides
            fd = (lp - lm)[same].sum() / (2 * H_VERTEX)
            gsub, _, _ = O.mirror_grad(V, F, o[same], d[same], pid[:, same], w[:, same])
            assert _fd_close(fd, gsub[vi, c]), (vi, c, fd, gsub[vi, c])
            checked += 1
    assert checked == 15
