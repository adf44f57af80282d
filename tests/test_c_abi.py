"""The batch API from a plain C program (tests/c_abi_search.c): include/tcme.h
compiles as C and the program links against libtcme.so (CPU); on a GPU it
searches every quadtree node of a frame and its results equal the oracle's."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "c_abi_search.c")
LIBDIR = os.path.join(ROOT, "tcodec_amd", "lib")


def build(tmp_path):
    exe = str(tmp_path / "c_abi_search")
    subprocess.run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
                    SRC, "-o", exe, f"-L{LIBDIR}", "-ltcme", f"-Wl,-rpath,{LIBDIR}"],
                   check=True, capture_output=True, text=True)
    return exe


def test_c_program_builds_against_the_header_and_library(tmp_path):
    if not os.path.exists(os.path.join(LIBDIR, "libtcme.so")):
        pytest.skip("libtcme.so not built")
    assert os.path.exists(build(tmp_path))


@pytest.mark.gpu
def test_c_program_searches_a_frame(tmp_path, orc):
    from tcodec_amd import synth, workload
    exe = build(tmp_path)
    fr = synth.scene_video(320, 192, 3, seed=5, rects=5, pan=(0.75, 0.25))
    ref, cur = fr[0][0], fr[2][0]
    nodes = workload.qt_nodes(320, 192)
    rows = np.stack([nodes[:, 0], nodes[:, 1], nodes[:, 0], nodes[:, 1], nodes[:, 2],
                     np.full(len(nodes), workload.SR_B)], 1).astype(np.int32)
    for name, arr in (("ref.y", ref), ("cur.y", cur), ("calls.i32", rows)):
        np.ascontiguousarray(arr).tofile(str(tmp_path / name))
    out = tmp_path / "out.i32"
    p = subprocess.run([exe, "320", "192", str(tmp_path / "ref.y"), str(tmp_path / "cur.y"),
                        str(tmp_path / "calls.i32"), str(len(rows)), str(out)],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr
    got = np.fromfile(str(out), np.int32).reshape(-1, 4)
    assert (got[:, 3] == 0).all()
    want = orc.motion_est_batch(ref, cur, rows)
    assert np.array_equal(got[:, :3], want)
