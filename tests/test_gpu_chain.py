"""The ME-only predictor chain on the GPU (tcme_search_chain) against the
oracle's chain, CTU by CTU and call by call: node order, centres (qt_mvp of
the chain's grid), references and results."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check(got, want, label):
    assert got.shape == want.shape, (label, got.shape, want.shape)
    bad = np.nonzero((got != want).any(1))[0]
    assert bad.size == 0, (label, f"{bad.size} rows differ; first row {bad[0]}: got "
                                  f"{got[bad[0]]} want {want[bad[0]]}")


@pytest.mark.parametrize("w,h,lam", [(192, 128, 4), (200, 136, 4), (320, 240, 0), (96, 80, 16)])
def test_chain_matches_oracle_small_frames(engine, orc, w, h, lam):
    from tcodec_amd import synth
    log(f"chain {w}x{h} lambda {lam}")
    fr = synth.scene_video(w, h, 4, seed=w + h, rects=5, pan=(2.0, -1.0))
    planes = [f[0] for f in fr]
    pics = [engine.pic(p) for p in planes]
    try:
        frames = [(pics[3], [pics[2], pics[1]], 24, lam),    # P with two references
                  (pics[2], [pics[1], pics[3]], 32, lam),    # B: forward, backward
                  (pics[1], [pics[0]], 16, lam)]             # one reference
        got = engine.chain(frames)
        for (cur, refs, sr, l), g, ci in zip(frames, got, (3, 2, 1)):
            ri = [pics.index(r) for r in refs]
            want = orc.chain_frame([planes[i] for i in ri], planes[ci], sr, l)
            check(g, want, f"{w}x{h} frame {ci}")
            # the first search of every CTU has a zero predictor: centre at the block
            first = np.r_[0, np.nonzero(np.diff(g[:, 0]))[0] + 1]
            assert (g[first, 4] == g[first, 2]).all() and (g[first, 5] == g[first, 3]).all()
    finally:
        for p in pics:
            engine.free(p)


def test_chain_full_hd_frame(engine, orc):
    """A 1920x1080 P frame of the 1080p scene generator against two
    references: every CTU's chain equal to the oracle's."""
    from tcodec_amd import synth
    log("chain 1080p")
    sc = synth.Scene(1920, 1080, 7, 48, (0.75, 0.25))
    planes = [sc.frame(t)[0] for t in range(3)]
    pics = [engine.pic(p) for p in planes]
    try:
        got = engine.chain([(pics[2], [pics[1], pics[0]], 24, 4)])[0]
        want = orc.chain_frame([planes[1], planes[0]], planes[2], 24, 4)
        check(got, want, "1080p")
        log(f"chain 1080p: {len(got)} searches")
    finally:
        for p in pics:
            engine.free(p)


def test_chain_rejects_bad_frames(engine):
    import torch
    from tcodec_amd import me
    a, b = engine.pic(np.zeros((64, 64), np.uint8)), engine.pic(np.zeros((64, 128), np.uint8))
    plane = engine.plane(np.zeros((64, 64), np.uint8))
    d_calls = torch.empty(8 * me.CHAIN_MAX_CALLS * 40, dtype=torch.uint8, device="cuda")
    d_counts = torch.empty(8, dtype=torch.int32, device="cuda")
    try:
        for cur, refs, sr, lam, cap in ((a, [b], 24, 4, 8),       # size mismatch
                                        (a, [plane], 24, 4, 8),   # reference without anchors
                                        (a, [], 24, 4, 8),        # no reference
                                        (a, [a], 24, -1, 8),      # negative lambda
                                        (a, [a], 24, 4, 0)):      # no room
            rec = np.zeros(1, me.CHAIN_FRAME_DTYPE)
            rec[0]["cur_pic"], rec[0]["nrefs"] = cur, len(refs)
            rec[0]["ref_pic"][:len(refs)] = refs
            rec[0]["sr"], rec[0]["lambda"] = sr, lam
            with pytest.raises(me.TcmeError):
                engine.chain_device(rec, d_calls.data_ptr(), d_counts.data_ptr(), cap)
    finally:
        for p in (a, b, plane):
            engine.free(p)
