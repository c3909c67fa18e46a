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
