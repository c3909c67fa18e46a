"""Builds libblazert_amd.so in-tree with hipcc for gfx950.

Sources: csrc/kernels.hip, csrc/mirror.hip, csrc/build.hip and csrc/distance.hip (device code), csrc/scene.cpp (host builder + C ABI) and
csrc/autotune.cpp (launch / leaf-size autotuner). Both are compiled with
-ffp-contract=off so that no multiply-add is fused (the results must round exactly like the reference's
separate multiply and add), and device code with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libblazert_amd.so")
ARCH = "gfx950"
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

COMMON = ["-O3", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", "-Wno-unused-function"]
DEVICE = ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-fno-gpu-flush-denormals-to-zero"]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[-1]}")
    return r


def build(verbose: bool = False, out: str = LIB, defines=()) -> str:
    """Compiles every source (always) and links libblazert_amd.so next to this file (or `out`)."""
    cc = hipcc()
    tmp = tempfile.mkdtemp(prefix="brt_build_")
    try:
        objs = []
        for src in ("kernels.hip", "mirror.hip", "build.hip", "distance.hip"):
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            _run([cc, *COMMON, *DEVICE, *[f"-D{d}" for d in defines], "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        for src in ("scene.cpp", "autotune.cpp"):
            obj = os.path.join(tmp, src.replace(".cpp", ".o"))
            _run([cc, *COMMON, "-x", "c++", "-pthread", "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include", "-c",
                  os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        out_tmp = out + ".tmp"
        _run([cc, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-pthread", *objs, "-o", out_tmp])
        os.replace(out_tmp, out)
        if verbose:
            print("built", out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


SANITIZED_LIB = os.path.join(HERE, "_asan", "libblazert_amd.so")


def build_sanitized(verbose: bool = False, out: str = SANITIZED_LIB) -> str:
    """The library with its host code (scene.cpp: builder, collections, C ABI; autotune.cpp) built by g++ with
    AddressSanitizer and UndefinedBehaviorSanitizer; the device code is compiled as usual. For CPU test runs
    (tests/run_sanitized.sh loads it through BLAZERT_AMD_LIB, with libasan / libubsan preloaded into Python)."""
    cc = hipcc()
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = tempfile.mkdtemp(prefix="brt_asan_")
    try:
        objs = []
        for src in ("kernels.hip", "mirror.hip", "build.hip", "distance.hip"):
            obj = os.path.join(tmp, src.replace(".hip", ".o"))
            _run([cc, *COMMON, *DEVICE, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        san = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-g"]
        for src in ("scene.cpp", "autotune.cpp"):
            obj = os.path.join(tmp, src.replace(".cpp", ".o"))
            _run(["g++", "-O1", "-std=c++17", "-fPIC", "-ffp-contract=off", "-pthread", "-D__HIP_PLATFORM_AMD__",
                  f"-I{ROCM}/include", *san, "-c", os.path.join(CSRC, src), "-o", obj])
            objs.append(obj)
        out_tmp = out + ".tmp"
        _run(["g++", "-shared", "-fPIC", "-pthread", *san, *objs, f"-L{ROCM}/lib", "-lamdhip64",
              f"-Wl,-rpath,{ROCM}/lib", "-o", out_tmp])
        os.replace(out_tmp, out)
        if verbose:
            print("built", out)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return out


INCLUDE = os.path.join(os.path.dirname(HERE), "include")
USER_GEOMETRY_EXAMPLE = os.path.join(os.path.dirname(HERE), "examples", "user_geometry", "user_geometry.hip")
USER_GEOMETRY_EXAMPLE_LIB = os.path.join(os.path.dirname(HERE), "examples", "user_geometry", "libuser_geometry.so")


def build_user_geometry(src: str, out: str, verbose: bool = False) -> str:
    """Compiles a user-geometry translation unit (device collections + include/blazert_amd/user_geometry.h) into a
    shared library linked against libblazert_amd.so, with the flags the exact path needs."""
    cc = hipcc()
    out_tmp = out + ".tmp"
    _run([cc, *COMMON, *DEVICE, f"-I{INCLUDE}", "-shared", src, f"-L{HERE}", "-lblazert_amd",
          f"-Wl,-rpath,{HERE}", "-o", out_tmp])
    os.replace(out_tmp, out)
    if verbose:
        print("built", out)
    return out


MICROBENCH_SRC = os.path.join(CSRC, "microbench.hip")
MICROBENCH_LIB = os.path.join(HERE, "libbrt_microbench.so")


def build_microbench(verbose: bool = False, out: str = MICROBENCH_LIB) -> str:
    """Compiles the ceiling microbenchmarks (csrc/microbench.hip: dependent random line chains, random line gathers,
    random FP64 atomics) into their own library, run by `bench.py` beside the kernels it measures."""
    cc = hipcc()
    out_tmp = out + ".tmp"
    _run([cc, *COMMON, *DEVICE, "-shared", MICROBENCH_SRC, "-o", out_tmp])
    os.replace(out_tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    if sys.argv[1:] == ["--sanitize"]:
        build_sanitized(verbose=True)
    else:
        build(verbose=True)
        build_user_geometry(USER_GEOMETRY_EXAMPLE, USER_GEOMETRY_EXAMPLE_LIB, verbose=True)
        build_microbench(verbose=True)
