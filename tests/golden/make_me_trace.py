"""Writes the golden ME call traces in tests/golden/ from the reference encoder.

Runs oracle/_ref/tcenc_hook (the reference encoder compiled from its own
sources, with every tc_motion_est call recorded by oracle/me_shim.c) on small
seeded clips from tcodec_amd.synth, and stores for each clip:

  ref_pocs, refs   recon luma planes the calls searched, keyed by POC
  cur_ids, curs    source luma planes (coding order)
  calls            int32 [N, 16]: cur_id px py ref_poc cx cy blk sr w h
                   mv_x mv_y sad orig_stride 0 0  (outputs from the reference)
  tcv_sha256       hash of the encoded stream

Usage: python tests/golden/make_me_trace.py   (needs oracle/_ref built)
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from tcodec_amd import synth  # noqa: E402

HOOK = os.path.join(ROOT, "oracle", "_ref", "tcenc_hook")

# name -> (generator, w, h, frames, encoder flags)
CLIPS = {
    "p_gradient_96x64": ("gradient", 96, 64, 4, ["--v2", "-q", "32"]),
    "p_multiref_128x96": ("scene", 128, 96, 5, ["--v2", "--profile", "1", "-p", "2", "-q", "27"]),
    "b_gop4_128x96": ("scene", 128, 96, 5, ["--v2", "--profile", "1", "-p", "2", "-q", "32", "-b"]),
}


def make_frames(kind, w, h, n, seed=5):
    if kind == "gradient":
        return synth.gradient_video(w, h, n, seed=seed)
    return synth.scene_video(w, h, n, seed=seed, rects=4, pan=(0.75, 0.25))


def parse_trace(path, w, h):
    raw = np.fromfile(path, dtype=np.uint8)
    pos, refs, curs, calls = 0, {}, {}, []
    while pos < raw.size:
        tag, a, b, c = np.frombuffer(raw[pos:pos + 16].tobytes(), "<i4")
        pos += 16
        if tag == 1:
            plane = raw[pos:pos + c].reshape(h, w).copy()
            pos += c
            (refs if a == 0 else curs)[int(b)] = plane
        elif tag == 2:
            calls.append(np.frombuffer(raw[pos:pos + 4 * a].tobytes(), "<i4").copy())
            pos += 4 * a
        else:
            raise ValueError(f"bad trace tag {tag}")
    return refs, curs, np.array(calls, np.int32)


def record(name):
    kind, w, h, n, flags = CLIPS[name]
    frames = make_frames(kind, w, h, n)
    with tempfile.TemporaryDirectory() as td:
        yuv = os.path.join(td, "in.yuv")
        synth.write_yuv(yuv, frames)
        tcv = os.path.join(td, "out.tcv")
        trace = os.path.join(td, "trace.bin")
        env = dict(os.environ, TCME_SHIM="trace", TCME_TRACE=trace)
        subprocess.run([HOOK, "-w", str(w), "-h", str(h), *flags, "-o", tcv, yuv], check=True,
                       env=env, capture_output=True)
        refs, curs, calls = parse_trace(trace, w, h)
        sha = hashlib.sha256(open(tcv, "rb").read()).hexdigest()
    rp = sorted(refs)
    cp = sorted(curs)
    return dict(ref_pocs=np.array(rp, np.int32), refs=np.stack([refs[p] for p in rp]),
                cur_ids=np.array(cp, np.int32), curs=np.stack([curs[i] for i in cp]),
                calls=calls, tcv_sha256=np.array(sha), width=np.int32(w), height=np.int32(h))


def main():
    for name in CLIPS:
        d = record(name)
        out = os.path.join(HERE, f"me_trace_{name}.npz")
        np.savez_compressed(out, **d)
        print(f"{out}: {d['calls'].shape[0]} calls, {len(d['ref_pocs'])} refs, "
              f"{os.path.getsize(out)} bytes")


if __name__ == "__main__":
    main()
