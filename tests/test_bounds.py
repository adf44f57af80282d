"""Descriptor bounds (tcodec_amd/csrc/tcme_bounds.h) at the int32 edges, on the
CPU: the same inline functions the kernels use, compiled for the host with
UBSan so that an overflowing sum fails the run."""
import os
import subprocess

from conftest import ROOT


def test_bounds_helpers_at_int32_edges(tmp_path):
    exe = str(tmp_path / "bounds_check")
    subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", "-fsanitize=undefined",
                    "-fno-sanitize-recover=all", "-I", os.path.join(ROOT, "tcodec_amd", "csrc"),
                    os.path.join(ROOT, "tests", "bounds_check.cpp"), "-o", exe],
                   check=True, capture_output=True, text=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
