"""The drop-in header's blaze arithmetic, used the way the reference's path tracer uses it
(examples/path_tracer/main.cc:103-126, 219-240, 255-270): tests/cpp/dropin_arith.cpp compiled against
include/blazert_amd/blazert.hpp, every value compared bit for bit with the same expressions evaluated here in IEEE
double / single precision (blaze's conventions: dot = (a0 b0 + a1 b1) + a2 b2, normalize = v * (1 / norm),
v / s = v * (1 / s), componentwise vector products, row-major matrix-vector rows summed left to right)."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_arith.cpp")
LIBDIR = os.path.join(ROOT, "blazert_amd")


def run(tmp_path, ft):
    exe = str(tmp_path / f"dropin_arith_{ft}")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", f"-DDROPIN_FT={ft}", "-I",
                    os.path.join(ROOT, "include"), SRC, "-L", LIBDIR, "-lblazert_amd", f"-Wl,-rpath,{LIBDIR}", "-o",
                    exe], check=True)
    out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout
    vals = {}
    for line in out.split("\n"):
        if line:
            k, v = line.split()
            vals[k] = float.fromhex(v) if ("x" in v or "." in v or "inf" in v) else int(v)
    return vals


def expected(F):
    """The caller's computation in scalar type F (float for double, np.float32 for float)."""
    sqrt = (lambda a: F(math.sqrt(a))) if F is float else (lambda a: np.sqrt(F(a)))
    V = lambda *a: [F(x) for x in a]
    add = lambda a, b: [a[k] + b[k] for k in range(3)]
    sub = lambda a, b: [a[k] - b[k] for k in range(3)]
    smul = lambda s, a: [s * a[k] for k in range(3)]
    vmul = lambda a, s: [a[k] * s for k in range(3)]
    dot = lambda a, b: a[0] * b[0] + a[1] * b[1] + a[2] * b[2]
    cross = lambda a, b: [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]
    length = lambda a: sqrt(dot(a, a))

    def normalize(a):
        n = length(a)
        return vmul(a, F(1) / n if n != 0 else F(0))

    def reflect(I, N):
        return sub(I, smul(F(2.0) * dot(I, N), N))

    def refract(I, N, eta):
        NdotI = dot(N, I)
        k = F(1.0) - eta * eta * (F(1.0) - NdotI * NdotI)
        if k < 0:
            return V(0, 0, 0)
        return sub(smul(eta, I), smul(eta * NdotI + sqrt(k), N))

    v0, v1, v2 = V(0.1, 0.2, 0.3), V(1.5, -0.25, 0.75), V(-0.5, 2.0, 1.25)
    x = V(0.3, -1.0, 4.0)
    Xi1, Xi2 = F(0.375), F(0.8125)
    Xi1_ = sqrt(Xi1)
    c0, c1, c2 = F(1.0) - Xi1_, Xi1_ * (F(1.0) - Xi2), Xi1_ * Xi2
    e = {}
    tmp = cross(sub(v1, v0), sub(v2, v0))
    norm = normalize(tmp)
    area = length(tmp) / F(2.0)
    lp = add(add(smul(c0, v0), smul(c1, v1)), smul(c2, v2))
    d = sub(lp, x)
    dist = length(d)
    d = normalize(d)
    cos_l = max(dot([-c for c in d], norm), F(0.0))
    rad = vmul(V(1.0, 0.5, 0.25), cos_l)
    refl = reflect(d, norm)
    refr = refract(d, norm, F(1.0 / 1.5))
    tir = refract(d, norm, F(4.0))
    r0 = F(0.2) * F(0.2)
    b = F(1.0) - dot(norm, d)
    fres = r0 + (F(1.0) - r0) * (b * b * b * b * b)
    acc = add(V(0.5, 0.5, 0.5), refl)
    acc = vmul(acc, F(3.0))
    acc = sub(acc, vmul(refr, F(1) / F(4.0)))
    schur = [refl[k] * refr[k] for k in range(3)]
    raw = sub(lp, x)
    e.update({f"tmp{k}": tmp[k] for k in range(3)})
    e.update({f"norm{k}": norm[k] for k in range(3)})
    e["area"] = area
    e.update({f"lp{k}": lp[k] for k in range(3)})
    e["dstDist"] = dist
    e.update({f"dstDir{k}": d[k] for k in range(3)})
    e["cosAtLight"] = cos_l
    for name, vec in (("rad", rad), ("refl", refl), ("refr", refr), ("tir", tir), ("acc", acc), ("schur", schur),
                      ("raydir", normalize(raw)), ("rayinv", [F(1) / c for c in raw])):
        e.update({f"{name}{k}": vec[k] for k in range(3)})
    e["F"] = fres
    e.update({f"raysign{k}": int(raw[k] < 0) for k in range(3)})
    e["raytmax"] = dist - F(0.00001)
    # trans(R) with R = [[0,-1,0],[1,0,0],[0,0,1]]: rows of R^T are R's columns
    Rt = [V(0.0, 1.0, 0.0), V(-1.0, 0.0, 0.0), V(0.0, 0.0, 1.0)]
    e.update({f"local{k}": dot(Rt[k], raw) for k in range(3)})
    return e


@pytest.mark.parametrize("ft,F", [("double", float), ("float", np.float32)])
def test_path_tracer_arithmetic_bit_exact(tmp_path, ft, F):
    got = run(tmp_path, ft)
    exp = expected(F)
    assert set(got) == set(exp)
    for k, v in exp.items():
        g = got[k]
        if isinstance(v, int):
            assert g == v, k
        else:
            assert np.float64(g).tobytes() == np.float64(v).tobytes(), (k, g, float(v))
    # the refraction past the critical angle is the zero vector, as in the reference caller
    assert got["tir0"] == got["tir1"] == got["tir2"] == 0.0
