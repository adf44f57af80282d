"""End to end: the reference encoder with the GPU backend linked in place of
its own tc_motion_est writes byte-identical .tcv, and the reference decoder
turns both into identical YUV.

Binaries (built by oracle/Makefile from the reference sources):
  oracle/_ref/tcenc         reference encoder, CPU motion estimation
  oracle/_ref/tcenc_gpu     the same objects linked with the frame-level hook
                            (tcodec_amd/glue/tcme_encoder_hook.c) and libtcme.so
  oracle/_ref/tcenc_gpu1    the same objects, tc_motion_est from libtcme.so
                            (the per-call drop-in)
  oracle/_ref/tcenc_fcheck  the frame-level link with every call also run by
                            the reference and compared (oracle/me_shim.c)
  oracle/_ref/tcdec         reference decoder

Each case runs three encoders at once (separate processes): the CPU encoder,
the frame-level GPU encoder (bytes compared), and the check-mode encoder,
which runs every call through the frame-level path and the reference and
compares them.  The check mode also reports calls per call site, so a case
can require that a site fired (multi-reference, global-motion centre, B
forward/backward).
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

# name: (generator, w, h, frames, generator options, encoder flags, sites that must fire)
CASES = {
    # the CPU-runnable configuration: gradient 320x240, 8 frames, v2 P, 1 ref, QP 32
    "config1_gradient_320x240_p_q32": ("gradient", 320, 240, 8, {}, ["--v2", "-q", "32"], ()),
    "scene_640x360_p_multiref_q27": ("scene", 640, 360, 4, {"rects": 8, "pan": (0.75, 0.25)},
                                     ["--v2", "--profile", "1", "-p", "2", "-q", "27"],
                                     ("slot1",)),
    # a pan of (8, 4) px per frame: the global estimate (4 px steps) is
    # 32 quarter samples, above the 16 that gates the global-centre search
    "scene_640x360_pan8_p_multiref_q27": ("scene", 640, 360, 4, {"rects": 8, "pan": (8.0, 4.0)},
                                          ["--v2", "--profile", "1", "-p", "2", "-q", "27"],
                                          ("slot1", "glob_centre")),
    "scene_480x272_bgop4_q32": ("scene", 480, 272, 9, {"rects": 8, "pan": (0.75, 0.25)},
                                ["--v2", "--profile", "1", "-p", "2", "-q", "32", "-b"],
                                ("bframe",)),
    # the 720p configuration in full: 30 frames, 16 rectangles, sigma 2,
    # P-only, 4-slot buffer (profile 1 searches dpb[0] and dpb[1]), QP 27
    "config2_scene_1280x720_30f_p_multiref_q27": (
        "scene", 1280, 720, 30, {"rects": 16, "pan": (0.0, 0.0)},
        ["--v2", "--profile", "1", "-p", "2", "-q", "27"], ("slot1",)),
    # the 1080p configuration's generator (48 rectangles, pan (0.75, 0.25)),
    # 17 frames: four GOP4 ladders of hierarchical B pictures, QP 32
    "config3_scene_1920x1080_17f_bgop4_q32": (
        "scene", 1920, 1080, 17, {"rects": 48, "pan": (0.75, 0.25)},
        ["--v2", "--profile", "1", "-p", "2", "-q", "32", "-b"], ("bframe",)),
    # v1 with 4 threads and no entropy coding: WPP rows on the encoder's
    # thread pool call tc_motion_est concurrently (8x8 stack blocks, stride 8)
    "gradient_320x240_v1_wpp_t4": ("gradient", 320, 240, 8, {}, ["-q", "32", "-t", "4"], ()),
}


def need(path):
    if not os.path.exists(path):
        pytest.skip(f"{path} not built")


def sha(path):
    return hashlib.sha256(open(path, "rb").read()).hexdigest()


@pytest.fixture(scope="module")
def clips(tmp_path_factory):
    from tcodec_amd import synth
    d = tmp_path_factory.mktemp("clips")
    out = {}
    for name, (kind, w, h, n, opts, flags, sites) in CASES.items():
        if kind == "gradient":
            frames = synth.gradient_video(w, h, n, seed=1)
        else:
            frames = synth.scene_video(w, h, n, seed=4, **opts)
        p = str(d / f"{name}.yuv")
        synth.write_yuv(p, frames)
        out[name] = (p, w, h, flags, sites)
    return out


def site_lines(stderr):
    out = []
    for ln in stderr.splitlines():
        if ln.startswith("TCME_SITE"):
            out.append({k: int(v, 0) for k, v in (kv.split("=") for kv in ln.split()[1:])})
    return out


@pytest.mark.timeout(900)
@pytest.mark.parametrize("name", list(CASES))
def test_dropin_encoder_bytes_identical_and_every_call_matches(clips, tmp_path, capfd, name):
    for b in ("tcenc", "tcenc_gpu", "tcdec", "tcenc_fcheck"):
        need(os.path.join(REF, b))
    yuv, w, h, flags, sites = clips[name]
    args = ["-w", str(w), "-h", str(h), *flags]
    out = {k: str(tmp_path / f"{k}.tcv") for k in ("cpu", "gpu", "chk")}
    # the check-mode encoder runs verbose (-v: one stderr line per coded
    # picture, no effect on the bytes), which the test relays as progress
    cmds = {"cpu": [os.path.join(REF, "tcenc"), *args, "-o", out["cpu"], yuv],
            "gpu": [os.path.join(REF, "tcenc_gpu"), *args, "-o", out["gpu"], yuv],
            "chk": [os.path.join(REF, "tcenc_fcheck"), *args, "-v", "-o", out["chk"], yuv]}
    env = {"cpu": None, "gpu": dict(os.environ, TCME_HOOK_STATS="1"),
           "chk": dict(os.environ, TCME_SHIM="check")}
    print(f"{name}: three encoders at once", flush=True)
    sys.stdout.flush()
    logs = {k: open(tmp_path / f"{k}.err", "w+") for k in ("cpu", "gpu")}
    procs = {k: subprocess.Popen(cmds[k], stdout=subprocess.DEVNULL, stderr=logs[k], text=True,
                                 env=env[k]) for k in ("cpu", "gpu")}
    procs["chk"] = subprocess.Popen(cmds["chk"], stdout=subprocess.DEVNULL,
                                    stderr=subprocess.PIPE, text=True, env=env["chk"])
    chk_err = []
    for ln in procs["chk"].stderr:
        chk_err.append(ln)
        if ln.startswith("Frame"):
            # the test's own output is captured; a coded picture is reported
            # on the terminal so that a long case shows its progress
            with capfd.disabled():
                print(f"{name}: check-mode {ln.strip()}", flush=True)
    res = {}
    for k, p in procs.items():
        p.wait(timeout=850)
        if k == "chk":
            se = "".join(chk_err)
        else:
            logs[k].seek(0)
            se = logs[k].read()
            logs[k].close()
        res[k] = (p.returncode, "", se)
        print(f"{name}: {k} exit {p.returncode}", flush=True)
    for k, (rc, so, se) in res.items():
        assert rc == 0, f"{k}: {se[-2000:]}"
    assert os.path.getsize(out["cpu"]) > 0
    assert sha(out["cpu"]) == sha(out["gpu"]), f"{name}: .tcv differs"
    assert sha(out["cpu"]) == sha(out["chk"])
    # every call of the check-mode encoder agreed
    se = res["chk"][2]
    line = [ln for ln in se.splitlines() if ln.startswith("TCME_SHIM")][-1]
    print(line)
    fields = dict(kv.split("=") for kv in line.split()[1:])
    assert int(fields["calls"]) > 0
    assert int(fields["mismatches"]) == 0, se[-2000:]
    st = site_lines(se)
    print("\n".join(str(s) for s in st))
    assert sum(s["calls"] for s in st) == int(fields["calls"])
    for key in sites:
        # a site whose every call is of this kind, with calls
        assert any(s["calls"] > 0 and s[key] == s["calls"] for s in st), (key, st)
    # the frame-level encoder's counters: every call answered one way
    fl = [ln for ln in res["gpu"][2].splitlines() if ln.startswith("TCME_FRAME")]
    assert fl, res["gpu"][2][-2000:]
    fs = {k: float(v) for k, v in (kv.split("=") for kv in fl[-1].split()[1:])}
    print(fl[-1])
    assert fs["calls"] == int(fields["calls"])
    assert fs["hits_spec"] + fs["hits_batch"] + fs["misses"] + fs["uncached"] == fs["calls"]
    if "--v2" in flags:
        assert fs["uncached"] == 0 and fs["hits_spec"] > 0
    # the decoder turns both streams into the same pictures
    ycpu, ygpu = str(tmp_path / "cpu.yuv"), str(tmp_path / "gpu.yuv")
    subprocess.run([os.path.join(REF, "tcdec"), out["cpu"], ycpu], check=True,
                   capture_output=True, timeout=600)
    subprocess.run([os.path.join(REF, "tcdec"), out["gpu"], ygpu], check=True,
                   capture_output=True, timeout=600)
    assert sha(ycpu) == sha(ygpu)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name", ["config1_gradient_320x240_p_q32", "gradient_320x240_v1_wpp_t4"])
def test_per_call_dropin_encoder_bytes_identical(clips, tmp_path, name):
    """The per-call link (libtcme.so's tc_motion_est in place of motion.c's,
    no hook): the same bytes as the CPU encoder."""
    for b in ("tcenc", "tcenc_gpu1"):
        need(os.path.join(REF, b))
    yuv, w, h, flags, sites = clips[name]
    args = ["-w", str(w), "-h", str(h), *flags]
    cpu, gpu = str(tmp_path / "cpu.tcv"), str(tmp_path / "gpu.tcv")
    subprocess.run([os.path.join(REF, "tcenc"), *args, "-o", cpu, yuv], check=True,
                   capture_output=True, timeout=280)
    subprocess.run([os.path.join(REF, "tcenc_gpu1"), *args, "-o", gpu, yuv], check=True,
                   capture_output=True, timeout=280)
    assert sha(cpu) == sha(gpu)
