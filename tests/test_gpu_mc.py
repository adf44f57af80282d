"""GPU motion compensation (tcme_predict_luma / tcme_predict_chroma) against
the oracle's tc_inter_predict / tc_inter_predict_chroma, block for block."""
import numpy as np
import pytest

from helpers import moving_pair

pytestmark = pytest.mark.gpu


def _preds(pic, mvs, sizes, stride=None):
    from tcodec_amd import me
    p = np.zeros(len(mvs), me.PRED_DTYPE)
    p["pic"] = pic
    p["mv_x"], p["mv_y"] = mvs[:, 0], mvs[:, 1]
    p["blk"] = sizes
    p["out_stride"] = sizes if stride is None else stride
    p["out_off"] = np.concatenate([[0], np.cumsum(sizes.astype(np.int64) ** 2)[:-1]]) \
        if stride is None else np.arange(len(mvs)) * stride * 64
    return p


@pytest.mark.parametrize("shape", [(61, 97), (128, 160), (40, 64), (72, 200)])
def test_gpu_luma_mc_matches_oracle(engine, orc, shape):
    h, w = shape
    print(f"luma mc {w}x{h}", flush=True)
    rng = np.random.default_rng(h * 31 + w)
    ref, _ = moving_pair(h, w, 5)
    rp = engine.pic(ref)
    sizes = rng.choice([1, 2, 4, 8, 16, 32, 64], 700)
    # positions over the frame and up to 40 samples beyond every edge, any phase
    mvs = np.stack([rng.integers(-160, 4 * w + 160, 700), rng.integers(-160, 4 * h + 160, 700)], 1)
    mvs[:60] &= ~3                                   # integer MVs too
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        got = out[o:o + n * n].reshape(n, n)
        want = orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(got, want), (k, mvs[k], n)
    engine.free(rp)


@pytest.mark.parametrize("n", [8, 16])
def test_gpu_luma_mc_runs_of_small_blocks(engine, orc, n):
    """Runs of 8x8 or 16x16 blocks (as a frame's call list holds them): every
    phase, integer MVs, runs whose blocks are all in the interior, runs with
    one block on or past the interior's edge (its slot then stages with
    clamped loads and fixes that block's border samples), and a rejected
    descriptor inside a run.  (The launch spreads a list over at least
    2,048 workgroups, so at these lengths a workgroup holds one or two
    blocks; test_gpu_luma_mc_frame_sized_list fills the slots.)"""
    from tcodec_amd import me, synth
    print(f"luma mc runs of {n}x{n} blocks", flush=True)
    h, w = 240, 320
    ref = synth.Scene(w, h, seed=12, rects=10).frame(1)[0]
    rp = engine.pic(ref)
    rng = np.random.default_rng(8 + n)
    for count in (4001, 4000, 6):
        sizes = np.full(count, n)
        fx = rng.integers(2, w - n - 2, count)
        fy = rng.integers(2, h - n - 2, count)
        mvs = np.stack([fx * 4 + rng.integers(0, 4, count), fy * 4 + rng.integers(0, 4, count)], 1)
        mvs[::7] &= ~3                                        # integer MVs
        if count > 100:
            # one block of some groups at the interior's edge or beyond it
            for k in range(1, count, 37):
                mvs[k] = [(w - n - 3 + int(rng.integers(0, 3))) * 4 + int(rng.integers(0, 4)),
                          int(rng.integers(-40, 8 * 4))]
        preds = _preds(rp, mvs, sizes)
        if count > 100:
            preds["pic"][50] = 999                            # rejected inside a group
        out, st = engine.predict(preds)
        for k in range(len(preds)):
            if count > 100 and k == 50:
                assert st[k] == -4
                continue
            assert st[k] == 0
            o = int(preds["out_off"][k])
            got = out[o:o + n * n].reshape(n, n)
            want = orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)
            assert np.array_equal(got, want), (count, k, mvs[k])
    engine.free(rp)


def test_gpu_luma_mc_interior_blocks(engine, orc):
    """Blocks whose every sample has its 6x6 apron inside the frame (the
    kernel's staged slots; sizes other than 8/16/32/64 sample by sample),
    each of the 16 phases, next to blocks on the edges of that region."""
    from tcodec_amd import synth
    print("luma mc interior 640x360", flush=True)
    h, w = 360, 640
    ref = synth.Scene(w, h, seed=11, rects=12).frame(0)[0]
    rp = engine.pic(ref)
    rng = np.random.default_rng(5)
    count = 3000
    sizes = rng.choice([8, 16, 32, 64, 12, 24, 48, 37, 50, 63], count)
    fx = np.array([rng.integers(2, w - n - 2) for n in sizes])
    fy = np.array([rng.integers(2, h - n - 2) for n in sizes])
    mvs = np.stack([fx * 4 + rng.integers(0, 4, count), fy * 4 + rng.integers(0, 4, count)], 1)
    mvs[:16] = np.stack([fx[:16] * 4 + np.arange(16) % 4, fy[:16] * 4 + np.arange(16) // 4], 1)
    # the last interior positions and one sample beyond them
    edge = count - 8
    mvs[edge:edge + 4, 0] = (w - sizes[edge:edge + 4] - 3 + np.arange(4) % 2) * 4 + 3
    mvs[edge + 4:, 1] = (h - sizes[edge + 4:] - 3 + np.arange(4) % 2) * 4 + 3
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    for k in range(count):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        want = orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(out[o:o + n * n].reshape(n, n), want), (k, mvs[k], n)
    engine.free(rp)


def test_gpu_luma_mc_far_outside_and_straddling(engine, orc):
    """Blocks of 8 .. 64 whose six-tap window leaves the picture: MVs up to
    2,000 samples past every edge and corner (every sample clamps to the
    border, far outside the apron the kernel stages), and blocks straddling
    each edge at every phase (six-tap, bilinear and nearest samples in one
    block)."""
    h, w = 90, 150
    print(f"luma mc far outside and straddling {w}x{h}", flush=True)
    rng = np.random.default_rng(41)
    ref, _ = moving_pair(h, w, 3)
    rp = engine.pic(ref)
    sizes, mvs = [], []
    for n in (8, 16, 32, 64):
        for dxq in (-2000, -300, -n - 7, -n - 2, -3, -1, 0, 1, 2, w - n - 2, w - n + 1, w + 4, w + 900):
            for dyq in (-1500, -n - 3, -2, 0, 3, h - n - 3, h - n + 2, h + 700):
                ph = int(rng.integers(0, 16))
                sizes.append(n)
                mvs.append((dxq * 4 + ph % 4, dyq * 4 + ph // 4))
    sizes, mvs = np.array(sizes), np.array(mvs)
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        got = out[o:o + n * n].reshape(n, n)
        want = orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(got, want), (k, mvs[k], n)
    engine.free(rp)


def test_gpu_luma_mc_every_position_and_phase(engine, orc):
    """1x1 predictions at every position of a 70x37 frame (two anchor tiles
    across, three down) at each of the 16 quarter-sample phases: an interior
    position reads the anchor fields (sample, hh, vv, cc) its phase selects,
    an edge one the bilinear or nearest ladder, so every anchor the plane
    holds is checked against the oracle's tc_inter_predict."""
    h, w = 37, 70
    print(f"luma mc every position and phase {w}x{h}", flush=True)
    ref = np.random.default_rng(23).integers(0, 256, (h, w), dtype=np.uint8)
    rp = engine.pic(ref)
    ys, xs, ph = np.meshgrid(np.arange(h), np.arange(w), np.arange(16), indexing="ij")
    mvs = np.stack([xs.ravel() * 4 + ph.ravel() % 4, ys.ravel() * 4 + ph.ravel() // 4], 1)
    preds = _preds(rp, mvs, np.ones(len(mvs), np.int64))
    out, st = engine.predict(preds)
    assert (st == 0).all()
    want = np.array([orc.inter_predict(ref, int(mx), int(my), 1)[0, 0] for mx, my in mvs])
    bad = np.nonzero(out[:len(mvs)] != want)[0]
    assert bad.size == 0, [(tuple(mvs[k]), int(out[k]), int(want[k])) for k in bad[:8]]
    engine.free(rp)


def test_gpu_uploads_back_to_back(engine):
    """Two uploads in a row from one host buffer that is rewritten in between,
    read through a row stride wider than the plane (and a device pitch wider
    still): each picture holds exactly the bytes of its own upload.  Read back
    by integer-MV chroma prediction, which returns the plane's samples with the
    coordinates clamped to the frame."""
    import ctypes

    from tcodec_amd import me
    rng = np.random.default_rng(17)
    h, w = 100, 136
    wide = rng.integers(0, 256, (h, 200), dtype=np.uint8)
    pa = engine.plane(np.zeros((h, w), np.uint8))
    pb = engine.plane(np.zeros((h, w), np.uint8))

    def upload(pid):  # the buffer itself, row stride 200 (no contiguous copy)
        assert engine.L.tcme_pic_upload(engine.h, pid, wide.ctypes.data_as(me.u8p), 200,
                                        ctypes.c_void_p(0)) == 0

    upload(pa)
    first = wide[:, :w].copy()
    wide[:] = 255 - wide
    upload(pb)
    second = wide[:, :w].copy()
    xs, ys = np.meshgrid(np.arange(-8, w + 8, 16), np.arange(-8, h + 8, 16))
    mvs = np.stack([xs.ravel() * 8, ys.ravel() * 8], 1)
    sizes = np.full(len(mvs), 16)
    for pid, want in ((pa, first), (pb, second)):
        out, st = engine.predict(_preds(pid, mvs, sizes), chroma=True)
        assert (st == 0).all()
        for k, (x, y) in enumerate(zip(xs.ravel(), ys.ravel())):
            rows = np.clip(np.arange(y, y + 16), 0, h - 1)
            cols = np.clip(np.arange(x, x + 16), 0, w - 1)
            got = out[k * 256:(k + 1) * 256].reshape(16, 16)
            assert np.array_equal(got, want[np.ix_(rows, cols)]), (pid, x, y)
    engine.free(pa)
    engine.free(pb)


@pytest.mark.parametrize("shape", [(30, 48), (64, 80), (20, 33)])
def test_gpu_chroma_mc_matches_oracle(engine, orc, shape):
    h, w = shape
    print(f"chroma mc {w}x{h}", flush=True)
    rng = np.random.default_rng(h * 7 + w)
    pl, _ = moving_pair(h, w, 9)
    cp = engine.plane(pl)
    sizes = rng.choice([1, 2, 4, 8, 16, 32], 600)
    mvs = np.stack([rng.integers(-200, 8 * w + 200, 600), rng.integers(-200, 8 * h + 200, 600)], 1)
    mvs[:50] &= ~7
    preds = _preds(cp, mvs, sizes)
    out, st = engine.predict(preds, chroma=True)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        got = out[o:o + n * n].reshape(n, n)
        want = orc.inter_predict_chroma(pl, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(got, want), (k, mvs[k], n)
    engine.free(cp)


def test_gpu_chroma_mc_small_blocks_near_edges(engine, orc):
    """Runs of 1x1..4x4 blocks between larger ones on a small plane (most of
    them clamped at an edge, so predicted one per wavefront), a rejected
    descriptor inside a run, and a list whose length is not a multiple of 4."""
    print("chroma mc small blocks near edges", flush=True)
    h, w = 40, 64
    pl, _ = moving_pair(h, w, 9)
    cp = engine.plane(pl)
    rng = np.random.default_rng(31)
    sizes = np.array([4] * 8 + [8] + [2, 3, 4, 1] + [16] + [4] * 5 + [1, 1, 3])
    mvs = np.stack([rng.integers(-80, 8 * w + 80, len(sizes)),
                    rng.integers(-80, 8 * h + 80, len(sizes))], 1)
    preds = _preds(cp, mvs, sizes)
    preds["pic"][2] = 999      # unknown picture, inside the first run of 4x4 blocks
    out, st = engine.predict(preds, chroma=True)
    assert st[2] == -4 and (np.delete(st, 2) == 0).all()
    for k in range(len(preds)):
        if k == 2:
            continue
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        want = orc.inter_predict_chroma(pl, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(out[o:o + n * n].reshape(n, n), want), (k, mvs[k], n)
    engine.free(cp)


def test_gpu_mc_strided_output_and_bad_descriptors(engine, orc):
    from tcodec_amd import me
    ref, _ = moving_pair(64, 96, 2)
    rp, pp = engine.pic(ref), engine.plane(ref)
    mvs = np.array([[37, 41], [100, 9], [8, 8], [5, 5], [5, 5], [5, 5]])
    p = _preds(rp, mvs, np.array([8, 16, 8, 8, 0, 65]), stride=80)
    p["pic"][2] = 999      # unknown picture
    p["pic"][3] = pp       # a plain plane cannot be a luma reference
    out, st = engine.predict(p)
    assert st.tolist() == [0, 0, -4, -4, -5, -5]
    for k in (0, 1):
        n = int(p["blk"][k])
        o = int(p["out_off"][k])
        got = np.stack([out[o + r * 80:o + r * 80 + n] for r in range(n)])
        assert np.array_equal(got, orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n))
    engine.free(rp)
    engine.free(pp)


def test_gpu_chroma_mc_interior_tasks(engine, orc):
    """Blocks of 4x4 .. 32x32 whose every sample's 2x2 neighbourhood is inside
    the plane (cut into 4x4 tasks shared by a workgroup's lanes), every
    1/8-sample fraction, in runs of one size and mixed as a CTU's quadtree
    levels are, next to blocks one sample past that region and other sizes,
    with a rejected descriptor among them."""
    from tcodec_amd import synth
    print("chroma mc interior tasks 320x240", flush=True)
    h, w = 240, 320
    pl = synth.Scene(w, h, seed=21, rects=9).frame(2)[1]
    pl = np.ascontiguousarray(np.repeat(np.repeat(pl, 2, 0), 2, 1)[:h, :w])
    cp = engine.plane(pl)
    rng = np.random.default_rng(77)
    sizes = np.concatenate([np.full(300, 4), np.full(70, 8), np.full(20, 16), np.full(9, 32),
                            np.tile([32, 16, 16, 16, 16, 8, 8, 8, 8, 4, 4, 4, 4], 40),
                            rng.choice([1, 2, 3, 5, 12, 20], 60)])
    count = len(sizes)
    xi = np.array([rng.integers(0, w - m) for m in sizes])
    yi = np.array([rng.integers(0, h - m) for m in sizes])
    mvs = np.stack([xi * 8 + rng.integers(0, 8, count), yi * 8 + rng.integers(0, 8, count)], 1)
    # every fraction on the first 64 blocks; some one sample past the region
    mvs[:64] = np.stack([xi[:64] * 8 + np.arange(64) % 8, yi[:64] * 8 + np.arange(64) // 8], 1)
    for k in range(5, count, 41):
        mvs[k, 0] = (w - sizes[k]) * 8 + int(rng.integers(0, 8))
    preds = _preds(cp, mvs, sizes)
    preds["pic"][100] = 999
    out, st = engine.predict(preds, chroma=True)
    assert st[100] == -4 and (np.delete(st, 100) == 0).all()
    for k in range(count):
        if k == 100:
            continue
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        want = orc.inter_predict_chroma(pl, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(out[o:o + n * n].reshape(n, n), want), (k, mvs[k], n)
    engine.free(cp)


def test_gpu_luma_mc_runs_of_one_phase(engine, orc):
    """Runs of 8x8 and 16x16 blocks sharing one quarter-sample phase, for each
    of the 16 phases: the kernel sorts a workgroup's blocks by the work their
    phase needs (dx even, dy zero) and gives a slot whose blocks all agree
    the task code of that kind, so these runs exercise every kind, interior
    blocks and blocks on the picture's edges alike (one or two blocks a
    workgroup at this length; the frame-sized list mixes kinds in full
    slots)."""
    from tcodec_amd import synth
    h, w = 144, 208
    print(f"luma mc runs of one phase {w}x{h}", flush=True)
    ref = synth.Scene(w, h, seed=31, rects=9).frame(2)[0]
    rp = engine.pic(ref)
    rng = np.random.default_rng(77)
    sizes, mvs = [], []
    for ph in range(16):
        for n, count in ((8, 96), (16, 40)):
            fx = rng.integers(-4, w - n + 4, count)
            fy = rng.integers(-4, h - n + 4, count)
            fx[: count // 2] = rng.integers(2, w - n - 3, count // 2)   # interior half
            fy[: count // 2] = rng.integers(2, h - n - 3, count // 2)
            sizes += [n] * count
            mvs += list(zip(fx * 4 + ph % 4, fy * 4 + ph // 4))
    sizes, mvs = np.array(sizes), np.array(mvs)
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        got = out[o:o + n * n].reshape(n, n)
        want = orc.inter_predict(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(got, want), (k, mvs[k], n)
    engine.free(rp)


@pytest.mark.parametrize("shape", [(1, 1), (2, 3), (7, 5), (9, 2), (4, 11), (13, 7), (3, 70)])
def test_gpu_mc_pictures_smaller_than_the_block(engine, orc, shape):
    """Pictures narrower or shorter than 4 samples and smaller than the
    block: every block reaches past the picture, every sample clamps
    (luma: 1..64 at every phase; chroma: 1..32 at every eighth)."""
    h, w = shape
    print(f"mc pictures smaller than the block {w}x{h}", flush=True)
    rng = np.random.default_rng(h * 100 + w)
    pl = rng.integers(0, 256, (h, w), dtype=np.uint8)
    rp, cp = engine.pic(pl), engine.plane(pl)
    sizes = rng.choice([1, 2, 4, 8, 12, 16, 32, 64], 300)
    mvs = np.stack([rng.integers(-4 * 80, 4 * (w + 12), 300), rng.integers(-4 * 80, 4 * (h + 12), 300)], 1)
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(sizes[k])
        o = int(preds["out_off"][k])
        want = orc.inter_predict(pl, int(mvs[k, 0]), int(mvs[k, 1]), n)
        assert np.array_equal(out[o:o + n * n].reshape(n, n), want), (k, mvs[k], n)
    csz = rng.choice([1, 2, 4, 8, 16, 32], 300)
    cmv = np.stack([rng.integers(-8 * 40, 8 * (w + 12), 300), rng.integers(-8 * 40, 8 * (h + 12), 300)], 1)
    preds = _preds(cp, cmv, csz)
    out, st = engine.predict(preds, chroma=True)
    assert (st == 0).all()
    for k in range(len(preds)):
        n = int(csz[k])
        o = int(preds["out_off"][k])
        want = orc.inter_predict_chroma(pl, int(cmv[k, 0]), int(cmv[k, 1]), n)
        assert np.array_equal(out[o:o + n * n].reshape(n, n), want), (k, cmv[k], n)
    engine.free(rp)
    engine.free(cp)


def _ctu_lists(rng, w, h, passes, reach, blk_scale=1):
    """Block lists as an encoder issues them: CTU after CTU, each CTU's
    quadtree nodes (1 64x64, 4 32x32, 16 16x16, 64 8x8, sizes divided by
    blk_scale), each pass over the picture with fresh MVs (block position
    plus up to +-reach samples, every fraction)."""
    sub = 8 // blk_scale  # fractions per sample: 4 luma, 8 chroma
    ctu = 64 // blk_scale
    nodes = [(n // blk_scale, bx, by) for n in (64, 32, 16, 8)
             for by in range(0, ctu, n // blk_scale) for bx in range(0, ctu, n // blk_scale)]
    ns, nx, ny = (np.array(v) for v in zip(*nodes))
    cx, cy = np.meshgrid(np.arange(0, w, ctu), np.arange(0, h, ctu))
    cx, cy = cx.ravel(), cy.ravel()
    sizes = np.tile(ns, len(cx) * passes)
    px = np.tile((cx[:, None] + nx[None, :]).ravel(), passes)
    py = np.tile((cy[:, None] + ny[None, :]).ravel(), passes)
    k = len(sizes)
    mvs = np.stack([(px + rng.integers(-reach, reach + 1, k)) * sub + rng.integers(0, sub, k),
                    (py + rng.integers(-reach, reach + 1, k)) * sub + rng.integers(0, sub, k)], 1)
    return sizes, mvs


def test_gpu_luma_mc_frame_sized_list(engine, orc):
    """A list as long as a 1080p frame's (142,800 blocks in CTU order over two
    references): the launch gives each of its 2,048 workgroups a chunk of
    ~70 descriptors, so its slots fill (16 8x8 blocks, 4 16x16, 32x32 and
    64x64 bands), blocks
    are sorted by kind with those reaching past the picture first, and runs
    of other sizes and a rejected descriptor sit among them.  Every block
    against the oracle."""
    from tcodec_amd import synth
    h, w = 720, 1280
    print(f"luma mc frame-sized list {w}x{h}", flush=True)
    sc = synth.Scene(w, h, seed=51, rects=16)
    refs = [sc.frame(2)[0], sc.frame(5)[0]]
    pics = [engine.pic(r) for r in refs]
    rng = np.random.default_rng(52)
    sizes, mvs = _ctu_lists(rng, w, h, 7, 24)
    which = (np.arange(len(sizes)) // (len(sizes) // 7)) % 2
    odd = rng.choice(len(sizes), 400, replace=False)
    sizes[odd] = rng.choice([4, 12, 24, 40], 400)
    preds = _preds(np.array(pics)[which], mvs, sizes)
    preds["pic"][777] = 999
    out, st = engine.predict(preds)
    assert st[777] == -4 and (np.delete(st, 777) == 0).all()
    want = np.zeros_like(out)
    for r in (0, 1):
        k = np.nonzero(which == r)[0]
        orc.inter_predict_list(refs[r], mvs[k], sizes[k], preds["out_off"][k], want)
    n2 = sizes.astype(np.int64) ** 2
    offs = preds["out_off"]
    want[offs[777]:offs[777] + n2[777]] = 0  # the rejected block is left untouched
    if not np.array_equal(out, want):
        bad = [k for k in range(len(sizes)) if not np.array_equal(
            out[offs[k]:offs[k] + n2[k]], want[offs[k]:offs[k] + n2[k]])]
        assert not bad, [(k, int(sizes[k]), tuple(mvs[k])) for k in bad[:8]]
    for p in pics:
        engine.free(p)


def test_gpu_chroma_mc_frame_sized_list(engine, orc):
    """A chroma list longer than 2,048 workgroups x 256 descriptors (652,800
    blocks of 4x4 .. 32x32 in CTU order over both planes): every workgroup
    takes two rounds of 256 descriptors, its 4x4 tasks numbered over all
    four wavefronts; blocks needing clamps, other sizes and a rejected
    descriptor among them.  Every block against the oracle."""
    from tcodec_amd import synth
    h, w = 720, 1280
    print(f"chroma mc frame-sized list {w // 2}x{h // 2}", flush=True)
    _, cb, cr = synth.Scene(w, h, seed=53, rects=16).frame(3)
    planes = [cb, cr]
    pics = [engine.plane(p) for p in planes]
    rng = np.random.default_rng(54)
    sizes, mvs = _ctu_lists(rng, w // 2, h // 2, 32, 20, blk_scale=2)
    which = (np.arange(len(sizes)) // 85) % 2
    odd = rng.choice(len(sizes), 300, replace=False)
    sizes[odd] = rng.choice([1, 2, 3, 6], 300)
    preds = _preds(np.array(pics)[which], mvs, sizes)
    preds["pic"][4321] = 999
    out, st = engine.predict(preds, chroma=True)
    assert st[4321] == -4 and (np.delete(st, 4321) == 0).all()
    want = np.zeros_like(out)
    for r in (0, 1):
        k = np.nonzero(which == r)[0]
        orc.inter_predict_list(planes[r], mvs[k], sizes[k], preds["out_off"][k], want, chroma=True)
    n2 = sizes.astype(np.int64) ** 2
    offs = preds["out_off"]
    want[offs[4321]:offs[4321] + n2[4321]] = 0  # the rejected block is left untouched
    if not np.array_equal(out, want):
        bad = [k for k in range(len(sizes)) if not np.array_equal(
            out[offs[k]:offs[k] + n2[k]], want[offs[k]:offs[k] + n2[k]])]
        assert not bad, [(k, int(sizes[k]), tuple(mvs[k])) for k in bad[:8]]
    for p in pics:
        engine.free(p)


@pytest.mark.parametrize("stride,shift", [(77, 0), (78, 1), (80, 2), (83, 3)])
def test_gpu_mc_unaligned_output(engine, orc, stride, shift):
    """Blocks written with a row stride and a start that are not multiples
    of 2 or 4 (the kernels then store byte by byte instead of 2 or 4
    samples at a time), interior and clamped blocks alike; the bytes between
    rows stay as they were."""
    h, w = 136, 200
    print(f"mc unaligned output stride {stride} shift {shift}", flush=True)
    rng = np.random.default_rng(stride * 4 + shift)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    rp, cp = engine.pic(ref), engine.plane(ref)
    for chroma, choices, sub in ((False, [8, 16, 32, 64, 12], 4), (True, [4, 8, 16, 32, 2], 8)):
        count = 6000
        sizes = rng.choice(choices, count)
        mvs = np.stack([rng.integers(-12 * sub, (w + 4) * sub, count),
                        rng.integers(-12 * sub, (h + 4) * sub, count)], 1)
        p = _preds(cp if chroma else rp, mvs, sizes, stride=stride)
        p["out_off"] = np.arange(count, dtype=np.int64) * stride * 64 + shift
        out, st = engine.predict(p, chroma=chroma)
        assert (st == 0).all()
        f = orc.inter_predict_chroma if chroma else orc.inter_predict
        gaps = np.ones(out.size, bool)
        for k in range(count):
            n, o = int(sizes[k]), int(p["out_off"][k])
            rows = o + stride * np.arange(n)[:, None] + np.arange(n)[None, :]
            assert np.array_equal(out[rows], f(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)), (chroma, k, n, tuple(mvs[k]))
            gaps[rows] = False
        assert not out[gaps].any()
    engine.free(rp)
    engine.free(cp)


def test_gpu_mc_extreme_motion_vectors(engine, orc):
    """MVs at the ends of the int32 range and around +-2^30 (every sample
    clamps to a corner or an edge of the picture), all sizes, luma and
    chroma."""
    h, w = 50, 70
    print("mc extreme motion vectors", flush=True)
    rng = np.random.default_rng(99)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    rp, cp = engine.pic(ref), engine.plane(ref)
    big = [-2 ** 31, -2 ** 31 + 5, -2 ** 30 - 3, -2 ** 20, 2 ** 20 + 1, 2 ** 30 + 2, 2 ** 31 - 1, 2 ** 31 - 6]
    small = [0, 9, 4 * w - 3]
    vals = big + small
    for chroma, choices in ((False, [1, 8, 16, 32, 64, 24]), (True, [1, 4, 8, 16, 32, 3])):
        mvs = np.array([(a, b) for a in vals for b in vals if a in big or b in big] * 2, np.int64)
        sizes = rng.choice(choices, len(mvs))
        p = _preds(cp if chroma else rp, mvs, sizes)
        out, st = engine.predict(p, chroma=chroma)
        assert (st == 0).all()
        f = orc.inter_predict_chroma if chroma else orc.inter_predict
        for k in range(len(mvs)):
            n, o = int(sizes[k]), int(p["out_off"][k])
            assert np.array_equal(out[o:o + n * n].reshape(n, n), f(ref, int(mvs[k, 0]), int(mvs[k, 1]), n)), \
                (chroma, k, n, tuple(mvs[k]))
    engine.free(rp)
    engine.free(cp)


def test_gpu_luma_mc_chunks_of_other_sizes(engine, orc):
    """A list long enough that every workgroup holds a full chunk of 128
    descriptors (270,000 blocks), nearly all of sizes other than 8/16/32/64
    (1..13), so a workgroup's slots of 128 samples for those blocks number
    more than 64 and are found by a binary search over all of them; every
    block against the oracle."""
    h, w = 144, 256
    print(f"luma mc chunks of other sizes {w}x{h}", flush=True)
    rng = np.random.default_rng(61)
    ref = rng.integers(0, 256, (h, w), dtype=np.uint8)
    rp = engine.pic(ref)
    count = 270_000
    sizes = rng.choice([1, 2, 3, 4, 5, 7, 12, 13, 8], count, p=[.1, .1, .1, .2, .1, .1, .1, .1, .1])
    mvs = np.stack([rng.integers(-40, 4 * w + 40, count), rng.integers(-40, 4 * h + 40, count)], 1)
    preds = _preds(rp, mvs, sizes)
    out, st = engine.predict(preds)
    assert (st == 0).all()
    want = orc.inter_predict_list(ref, mvs, sizes, preds["out_off"], np.zeros_like(out))
    if not np.array_equal(out, want):
        n2 = sizes.astype(np.int64) ** 2
        offs = preds["out_off"]
        bad = [k for k in range(count) if not np.array_equal(out[offs[k]:offs[k] + n2[k]],
                                                              want[offs[k]:offs[k] + n2[k]])]
        assert not bad, [(k, int(sizes[k]), tuple(mvs[k])) for k in bad[:8]]
    engine.free(rp)
