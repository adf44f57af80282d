"""End to end: the reference encoder with the GPU backend linked in place of
its own tc_motion_est writes byte-identical .tcv, and the reference decoder
turns both into identical YUV.

Binaries (built by oracle/Makefile from the reference sources):
  oracle/_ref/tcenc       reference encoder, CPU motion estimation
  oracle/_ref/tcenc_gpu   the same objects, tc_motion_est from libtcme.so
  oracle/_ref/tcenc_hook  the same objects, ME calls routed by oracle/me_shim.c
  oracle/_ref/tcdec       reference decoder
"""
import hashlib
import os
import subprocess
import sys

import pytest

from conftest import ROOT

pytestmark = pytest.mark.gpu

REF = os.path.join(ROOT, "oracle", "_ref")
LIB = os.path.join(ROOT, "tcodec_amd", "lib", "libtcme.so")

CASES = {
    # name: (generator, w, h, frames, flags)
    "gradient_320x240_p_q32": ("gradient", 320, 240, 8, ["--v2", "-q", "32"]),
    "scene_640x360_p_multiref_q27": ("scene", 640, 360, 4,
                                     ["--v2", "--profile", "1", "-p", "2", "-q", "27"]),
    "scene_480x272_bgop4_q32": ("scene", 480, 272, 9,
                                ["--v2", "--profile", "1", "-p", "2", "-q", "32", "-b"]),
    # the 720p and 1080p configurations' sizes and flags, on short clips
    "scene_1280x720_p_multiref_q27": ("scene", 1280, 720, 4,
                                      ["--v2", "--profile", "1", "-p", "2", "-q", "27"]),
    "scene_1920x1080_bgop4_q32": ("scene", 1920, 1080, 5,
                                  ["--v2", "--profile", "1", "-p", "2", "-q", "32", "-b"]),
}


def need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


def run(cmd, env=None):
    print("run:", " ".join(cmd), flush=True)
    sys.stdout.flush()
    return subprocess.run(cmd, check=True, capture_output=True, text=True, env=env,
                          timeout=100)


@pytest.fixture(scope="module")
def clips(tmp_path_factory):
    from tcodec_amd import synth
    d = tmp_path_factory.mktemp("clips")
    out = {}
    for name, (kind, w, h, n, flags) in CASES.items():
        if kind == "gradient":
            frames = synth.gradient_video(w, h, n, seed=1)
        else:
            frames = synth.scene_video(w, h, n, seed=4, rects=8, pan=(0.75, 0.25))
        p = str(d / f"{name}.yuv")
        synth.write_yuv(p, frames)
        out[name] = (p, w, h, flags)
    return out


@pytest.mark.parametrize("name", list(CASES))
def test_dropin_encoder_bytes_identical(clips, tmp_path, name):
    for b in ("tcenc", "tcenc_gpu", "tcdec"):
        need(os.path.join(REF, b))
    yuv, w, h, flags = clips[name]
    cpu, gpu = str(tmp_path / "cpu.tcv"), str(tmp_path / "gpu.tcv")
    args = ["-w", str(w), "-h", str(h), *flags]
    run([os.path.join(REF, "tcenc"), *args, "-o", cpu, yuv])
    run([os.path.join(REF, "tcenc_gpu"), *args, "-o", gpu, yuv])
    assert os.path.getsize(cpu) > 0
    assert sha(cpu) == sha(gpu), f"{name}: .tcv differs"
    ycpu, ygpu = str(tmp_path / "cpu.yuv"), str(tmp_path / "gpu.yuv")
    run([os.path.join(REF, "tcdec"), cpu, ycpu])
    run([os.path.join(REF, "tcdec"), gpu, ygpu])
    assert sha(ycpu) == sha(ygpu)


@pytest.mark.parametrize("name", list(CASES))
def test_every_encoder_call_matches(clips, tmp_path, name):
    """Check mode: each tc_motion_est the encoder makes runs on both the
    reference and the GPU; every {x, y, sad} must agree."""
    need(os.path.join(REF, "tcenc_hook"))
    yuv, w, h, flags = clips[name]
    env = dict(os.environ, TCME_SHIM="check", TCME_LIB=LIB)
    r = run([os.path.join(REF, "tcenc_hook"), "-w", str(w), "-h", str(h), *flags, "-o",
             str(tmp_path / "chk.tcv"), yuv], env=env)
    line = [ln for ln in r.stderr.splitlines() if ln.startswith("TCME_SHIM")][-1]
    print(line)
    fields = dict(kv.split("=") for kv in line.split()[1:])
    assert int(fields["calls"]) > 0
    assert int(fields["mismatches"]) == 0, r.stderr[-2000:]
