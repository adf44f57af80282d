"""GPU global-motion estimate and block SADs against the oracle and brute force."""
import numpy as np
import pytest

from helpers import moving_pair

pytestmark = pytest.mark.gpu


def test_gpu_global_mv_matches_oracle(engine, orc):
    from tcodec_amd import synth
    print("global mv", flush=True)
    rng = np.random.default_rng(8)
    planes, pairs, want = [], [], []
    for (h, w) in [(61, 97), (240, 320), (72, 200), (16, 16), (144, 176), (360, 640)]:
        for k in range(3):
            shift = tuple(int(v) for v in rng.integers(-40, 41, 2))
            ref, cur = moving_pair(h, w, 31 * k + w, shift=shift)
            rp, cp = engine.pic(ref), engine.pic(cur)
            planes += [rp, cp]
            pairs.append((cp, rp))
            want.append(orc.global_mv(cur, ref))
    sc = synth.Scene(640, 360, seed=4, rects=16, pan=(6.0, -3.0))
    f0, f3 = sc.frame(0)[0], sc.frame(3)[0]
    rp, cp = engine.pic(f0), engine.pic(f3)
    planes += [rp, cp]
    pairs.append((cp, rp))
    want.append(orc.global_mv(f3, f0))
    # A long ramp (triangle wave, period 1024 px, shifted 250 px) on a wide
    # picture: the walk travels 224 px, past the +-128 px displacement table,
    # so candidates beyond it are priced directly.
    x = np.arange(4096)
    f = np.mod(x, 1024) / 1024
    tri = lambda v: np.round(np.where(v < 0.5, v * 2, 2 - v * 2) * 255).astype(np.uint8)
    g = np.mod(x + 250, 1024) / 1024
    cur = np.ascontiguousarray(np.tile(tri(f), (256, 1)))
    ref = np.ascontiguousarray(np.tile(tri(g), (256, 1)))
    rp, cp = engine.pic(ref), engine.pic(cur)
    planes += [rp, cp]
    pairs.append((cp, rp))
    want.append(orc.global_mv(cur, ref))
    assert want[-1] == (-768, 0)
    got = engine.global_mv(pairs)
    assert (got[:, 2] == 0).all()
    assert [tuple(g[:2]) for g in got] == [tuple(x) for x in want]
    # planes[0] is 97x61, planes[6] 320x240: sizes differ; then an unknown id
    bad = engine.global_mv([(planes[0], planes[6]), (9999, planes[0])])
    assert bad[:, 2].tolist() == [-1, -4]
    for p in planes:
        engine.free(p)


def test_gpu_block_sads(engine):
    from tcodec_amd import me
    print("block sads", flush=True)
    rng = np.random.default_rng(12)
    h, w = 100, 150
    a, b = moving_pair(h, w, 3, shift=(2, -3))
    pa, pb = engine.pic(a), engine.plane(b)
    buf = rng.integers(0, 256, 64 * 64 * 50).astype(np.uint8)
    n = 400
    d = np.zeros(n, me.SAD_DTYPE)
    d["pic"] = pa
    d["w"] = rng.integers(1, 65, n)
    d["h"] = rng.integers(1, 65, n)
    d["x"] = [int(rng.integers(0, w - x + 1)) for x in d["w"]]
    d["y"] = [int(rng.integers(0, h - y + 1)) for y in d["h"]]
    use_pic = rng.random(n) < 0.5
    d["other_pic"] = np.where(use_pic, pb, -1)
    d["ox"] = [int(rng.integers(0, w - x + 1)) for x in d["w"]]
    d["oy"] = [int(rng.integers(0, h - y + 1)) for y in d["h"]]
    d["buf_stride"] = 64
    d["buf_off"] = np.arange(n) % 49 * 64 * 64
    got = engine.sad(d, buf)
    for k in range(n):
        bw, bh, x, y = int(d["w"][k]), int(d["h"][k]), int(d["x"][k]), int(d["y"][k])
        blk = a[y:y + bh, x:x + bw].astype(int)
        if use_pic[k]:
            ox, oy = int(d["ox"][k]), int(d["oy"][k])
            other = b[oy:oy + bh, ox:ox + bw].astype(int)
        else:
            o = int(d["buf_off"][k])
            other = np.stack([buf[o + r * 64:o + r * 64 + bw] for r in range(bh)]).astype(int)
        assert got[k] == np.abs(blk - other).sum(), k
    # co-located stability SADs over the CTU grid, clipped at the frame edge
    ctus = [(x, y, min(64, w - x), min(64, h - y)) for y in range(0, h, 64) for x in range(0, w, 64)]
    s = np.zeros(len(ctus), me.SAD_DTYPE)
    for k, (x, y, cw, ch) in enumerate(ctus):
        s[k] = (pa, x, y, cw, ch, pb, x, y, 0, 0, 0)
    got = engine.sad(s)
    want = [np.abs(a[y:y + ch, x:x + cw].astype(int) - b[y:y + ch, x:x + cw]).sum()
            for (x, y, cw, ch) in ctus]
    assert got.tolist() == want
    # rejected descriptors
    bad = np.zeros(4, me.SAD_DTYPE)
    bad[:] = (pa, 0, 0, 8, 8, pb, 0, 0, 0, 8, 0)
    bad["pic"][0] = 999
    bad["w"][1] = 65
    bad["x"][2] = w - 4
    bad["other_pic"][3] = -1      # buffer mode without a buffer
    assert engine.sad(bad).tolist() == [-4, -5, -1, -1]
    engine.free(pa)
    engine.free(pb)
