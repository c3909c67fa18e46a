"""The two order-dependent steps of the device builder, checked on the CPU against their sequential forms
(tests/cpp/build_premises.cpp): libstdc++'s std::partition equals the rank pairing the device partition uses, and
the lane-parallel SAH sweep selects the same bin as the sequential loop (empty bins and ties included)."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "build_premises.cpp")


def test_partition_pairing_and_sweep_selection(tmp_path):
    exe = str(tmp_path / "build_premises")
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", SRC, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "partition: rank pairing equals std::partition" in r.stdout
    assert "sweep: lane form equals the sequential loop" in r.stdout
