"""The C++17 drop-in header (include/blazert_amd/blazert.hpp) used like blazeRT's own scene API."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "dropin_scene.cpp")
LIBDIR = os.path.join(ROOT, "blazert_amd")


@pytest.fixture(scope="module")
def exe(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("cpp") / "dropin_scene")
    subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), SRC, "-L", LIBDIR,
                    "-lblazert_amd", f"-Wl,-rpath,{LIBDIR}", "-o", out], check=True)
    return out


def test_header_semantics_without_gpu(exe):
    r = subprocess.run([exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["0", "4294967295", "1", "4294967295", "1"]


@pytest.mark.gpu
def test_header_intersect1_renders_golden(exe):
    step = 20
    r = subprocess.run([exe, "render", "1080", "720", str(step)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.array([list(map(int, line.split())) for line in r.stdout.strip().splitlines()])
    labels = np.load(os.path.join(ROOT, "tests", "golden", "scene_primitives_labels.npz"))["labels"]
    expect = labels[got[:, 1] * 1080 + got[:, 0]]
    assert np.array_equal(got[:, 2], expect)
    assert (got[:, 2] > 0).sum() > 20


@pytest.fixture(scope="module")
def exe_f32(tmp_path_factory):
    """The same program with blazert::Scene<float> (-DDROPIN_FT=float)."""
    out = str(tmp_path_factory.mktemp("cppf") / "dropin_scene_f32")
    subprocess.run(["g++", "-std=c++17", "-O1", "-DDROPIN_FT=float", "-I", os.path.join(ROOT, "include"), SRC, "-L",
                    LIBDIR, "-lblazert_amd", f"-Wl,-rpath,{LIBDIR}", "-o", out], check=True)
    return out


def test_header_semantics_f32_without_gpu(exe_f32):
    r = subprocess.run([exe_f32], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.split() == ["0", "4294967295", "1", "4294967295", "1"]


@pytest.mark.gpu
def test_header_f32_render_matches_f32_oracle(exe_f32, oracle_mod):
    """Scene<float> through the C++ header against the FP32 oracle, rays built exactly as the program builds them
    ((x / float(W)) - 0.5 is evaluated in double, then narrowed)."""
    from blazert_amd import workloads as W
    step, w, h = 20, 1080, 720
    r = subprocess.run([exe_f32, "render", str(w), str(h), str(step)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    got = np.array([list(map(int, line.split())) for line in r.stdout.strip().splitlines()])
    xs, ys = got[:, 0].astype(np.float32), got[:, 1].astype(np.float32)
    dx = ((xs / np.float32(w)).astype(np.float64) - 0.5).astype(np.float32)
    dy = ((ys / np.float32(h)).astype(np.float64) - 0.5).astype(np.float32)
    dirs = np.stack([dx, dy, -np.ones_like(dx)], axis=1)
    org = np.broadcast_to(np.array([0.0, 5.0, 20.0], np.float32), dirs.shape)
    sp = W.scene_primitives()
    o = oracle_mod.SceneF32()
    o.add_cylinders(*sp["cylinders"])
    o.add_spheres(*sp["spheres"])
    o.commit()
    ref = o.intersect_batch(org, dirs)
    expect = np.where(ref["prim_id"] == 0xFFFFFFFF, 0, np.where(ref["prim_id"] == 0, 1, 2))
    assert np.array_equal(got[:, 2], expect)
    assert (expect > 0).sum() > 20
