"""GPU parity: the gfx950 search against the oracle, call for call, bit-exact.

Every test calls through the C ABI of libtcme.so (batch API via ctypes, or the
drop-in tc_motion_est symbol)."""
import sys

import numpy as np
import pytest

from helpers import golden_traces, moving_pair, random_calls

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_batch(orc, ref, cur, calls, stats=False):
    arr = np.stack([calls["px"], calls["py"], calls["cx"], calls["cy"], calls["blk"],
                    calls["sr"]], 1)
    return orc.motion_est_batch(ref, cur, arr, want_stats=stats)


def compare(res, want, calls, label):
    got = np.stack([res["mv_x"], res["mv_y"], res["sad"]], 1)
    bad = np.nonzero((got != want).any(1) | (res["status"] != 0))[0]
    assert bad.size == 0, (f"{label}: {bad.size}/{len(calls)} differ; first call "
                           f"{calls[bad[0]]} got {got[bad[0]]} want {want[bad[0]]} "
                           f"status {res['status'][bad[0]]}")


@pytest.mark.parametrize("trace", golden_traces(), ids=lambda t: t[0])
def test_gpu_replays_golden_trace(engine, trace):
    from tcodec_amd import me
    name, refs, curs, calls = trace
    log(f"golden {name}: {len(calls)} calls")
    rp = {p: engine.pic(pl) for p, pl in refs.items()}
    cp = {i: engine.pic(pl) for i, pl in curs.items()}
    try:
        c = me.make_calls([rp[int(v)] for v in calls[:, 3]], [cp[int(v)] for v in calls[:, 0]],
                          calls[:, 1], calls[:, 2], calls[:, 4], calls[:, 5], calls[:, 6],
                          calls[:, 7])
        res = engine.search(c)
        compare(res, calls[:, 10:13], c, name)
    finally:
        for p in list(rp.values()) + list(cp.values()):
            engine.free(p)


@pytest.mark.parametrize("shape", [(96, 128), (64, 80), (72, 200), (128, 64)])
def test_gpu_random_calls_vs_oracle(engine, orc, shape):
    from tcodec_amd import me
    h, w = shape
    rng = np.random.default_rng(h * 1000 + w)
    log(f"random calls {w}x{h}")
    for trial in range(4):
        ref, cur = moving_pair(h, w, trial + h, shift=tuple(int(v) for v in rng.integers(-9, 10, 2)))
        rp, cp = engine.pic(ref), engine.pic(cur)
        arr = random_calls(rng, w, h, 600)
        calls = me.make_calls(rp, cp, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], arr[:, 4],
                              arr[:, 5])
        res, st = engine.search(calls, stats=True)
        want, wst = oracle_batch(orc, ref, cur, calls, stats=True)
        compare(res, want, calls, f"{w}x{h} trial {trial}")
        # candidate counts follow the reference's sequential order exactly
        assert np.array_equal(st["int_evals"], wst[:, 0])
        assert np.array_equal(st["int_oob"], wst[:, 1])
        assert np.array_equal(st["q6_evals"], wst[:, 2])
        assert np.array_equal(st["qbil_evals"], wst[:, 3])
        engine.free(rp)
        engine.free(cp)


def test_gpu_every_node_of_a_synthetic_frame(engine, orc):
    """The throughput call set on a scene frame pair: all quadtree nodes,
    both ranges, with a moving texture so the walk travels."""
    from tcodec_amd import synth, workload
    log("scene 320x240 all nodes")
    fr = synth.scene_video(320, 240, 3, seed=3, rects=8, pan=(0.75, 0.25))
    ref, cur = fr[0][0], fr[2][0]
    rp, cp = engine.pic(ref), engine.pic(cur)
    nodes = workload.qt_nodes(320, 240)
    for sr in (workload.SR_P, workload.SR_B):
        calls = workload.frame_calls(nodes, [rp], cp, sr)
        res = engine.search(calls)
        compare(res, oracle_batch(orc, ref, cur, calls), calls, f"sr {sr}")
    engine.free(rp)
    engine.free(cp)


def test_gpu_edge_frames(engine, orc):
    """Frames at and below the block size and the 6-tap margins; centres far
    outside; the bilinear and skipped quarter-pel paths."""
    from tcodec_amd import me
    rng = np.random.default_rng(77)
    log("edge frames")
    for (h, w) in [(8, 8), (9, 13), (16, 12), (20, 40), (33, 17), (64, 64), (66, 70)]:
        ref = rng.integers(0, 256, (h, w)).astype(np.uint8)
        ref[:, : w // 2] = np.clip(ref[:, : w // 2] // 4 + 60, 0, 255)
        cur = rng.integers(0, 256, (64, 64)).astype(np.uint8)
        cur[8:40, 8:40] = ref[: min(h, 32), : min(w, 32)].mean()
        rp, cp = engine.pic(ref), engine.pic(cur)
        arr = []
        for n in (8, 16, 32, 64):
            for cx, cy in [(0, 0), (-5, 3), (w, h), (2, 1), (-100, 300)]:
                for sr in (1, 8, 24, 32, 64):
                    arr.append((0, 0, cx, cy, n, sr))
        arr = np.array(arr, np.int32)
        calls = me.make_calls(rp, cp, arr[:, 0], arr[:, 1], arr[:, 2], arr[:, 3], arr[:, 4],
                              arr[:, 5])
        res = engine.search(calls)
        want = oracle_batch(orc, ref, cur, calls)
        compare(res, want, calls, f"edge {w}x{h}")
        engine.free(rp)
        engine.free(cp)


def test_gpu_rejects_bad_tiling(engine):
    from tcodec_amd import me
    with pytest.raises(me.TcmeError):
        engine.set_tiling(2, 1, 2, 4)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 3, 4)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 1)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 4, 3, 85)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 4, 8, 0)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 4, 4, 97)
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 4, 0, 85, 1)    # tile_big without tile workgroups
    with pytest.raises(me.TcmeError):
        engine.set_tiling(1, 1, 2, 4, 8, 85, 2)


def test_gpu_rejects_bad_descriptors(engine):
    from tcodec_amd import me
    ref = np.zeros((32, 32), np.uint8)
    rp = engine.pic(ref)
    calls = me.make_calls([rp, rp, 999, rp, rp], [rp, rp, rp, rp, -1],
                          [0, 24, 0, 0, 0], [0, 0, 0, 0, 0], 0, 0, [8, 16, 8, 12, 8], 24)
    res = engine.search(calls)
    assert res["status"].tolist() == [0, -1, -4, -5, -4]
    engine.free(rp)


def test_gpu_int32_edge_descriptors(engine, orc):
    """Descriptors at the int32 edges: block origins near INT_MAX are rejected
    (their x + n used to wrap), and search ranges near INT_MAX give the
    oracle's answer (every large-hexagon point off the picture) in every
    launch shape and through the drop-in symbol."""
    from tcodec_amd import me
    log("int32 edges")
    big = 2**31 - 1
    ref, cur = moving_pair(96, 128, 5, shift=(2, -3))
    rp, cp = engine.pic(ref), engine.pic(cur)
    try:
        bad = me.make_calls(rp, cp, [big - 3, 0, big - 63, 0], [0, big - 3, 0, big], 0, 0,
                            [8, 8, 64, 16], 24)
        res = engine.search(bad)
        assert (res["status"] == -1).all(), res["status"]
        rng = np.random.default_rng(9)
        rows = random_calls(rng, 128, 96, 200, srs=(big, big - 3, 2**30, 4 * 128 + 7))
        rows[::7, 2] = big                       # raw centres at the edges too
        rows[3::7, 3] = -big - 1
        calls = me.make_calls(rp, cp, rows[:, 0], rows[:, 1], rows[:, 2], rows[:, 3],
                              rows[:, 4], rows[:, 5])
        want = orc.motion_est_batch(ref, cur, rows)
        t0 = engine.tiling()
        for t in ((1, 1, 1, 4, 8, 85, 1), (1, 1, 2, 4, 4, 85, 0), (1, 1, 1, 2, 0, 85, 0)):
            engine.set_tiling(*t)
            compare(engine.search(calls), want, calls, f"huge sr {t}")
        engine.set_tiling(*[getattr(t0, k) for k, _ in me.Tiling._fields_])
        for k in range(0, len(rows), 10):
            px, py, cx, cy, n, sr = (int(v) for v in rows[k])
            (mx, my), sad = me.motion_est(ref, 128, 128, 96, cur, py * 128 + px, 128, cx, cy,
                                          n, sr)
            assert (mx, my, sad) == tuple(want[k]), rows[k]
    finally:
        engine.free(rp)
        engine.free(cp)


def test_gpu_full_hd_sample(engine, orc):
    """A sample of the 1080p B workload checked exactly (the oracle is too
    slow for every call), plus the whole frame's call set run for status and
    a size-independent property: every result is at least as good as its
    clipped centre and lies on the frame."""
    from tcodec_amd import synth, workload
    log("1080p sample")
    sc = synth.Scene(1920, 1080, 7, 48, (0.75, 0.25))
    ref, cur = sc.frame(0)[0], sc.frame(2)[0]
    rp, cp = engine.pic(ref), engine.pic(cur)
    nodes = workload.qt_nodes(1920, 1080)
    calls = workload.frame_calls(nodes, [rp], cp, workload.SR_B)
    res = engine.search(calls)
    assert (res["status"] == 0).all()
    rng = np.random.default_rng(1)
    pick = np.sort(rng.choice(len(calls), 3000, replace=False))
    want = oracle_batch(orc, ref, cur, calls[pick])
    compare(res[pick], want, calls[pick], "1080p sample")
    n = calls["blk"]
    assert ((res["mv_x"] >> 2) >= -1).all() and ((res["mv_x"] >> 2) <= 1920 - n).all()
    assert ((res["mv_y"] >> 2) >= -1).all() and ((res["mv_y"] >> 2) <= 1080 - n).all()
    engine.free(rp)
    engine.free(cp)


def test_dropin_symbol_matches_oracle(orc):
    """tc_motion_est through the drop-in symbol: strided planes, v1-style
    8x8 stack blocks (stride 8), and a reference buffer reused with new
    content between calls (the mirror must notice)."""
    from tcodec_amd import me
    rng = np.random.default_rng(21)
    log("drop-in")
    h, w, stride = 96, 120, 128
    for trial in range(3):
        ref, cur = moving_pair(h, w, trial, shift=(2 - trial, 3))
        rbuf = np.zeros((h, stride), np.uint8)
        rbuf[:, :w] = ref
        for px, py, cx, cy, n, sr in random_calls(rng, w, h, 80):
            (mx, my), sad = me.motion_est(rbuf, stride, w, h, cur, py * w + px, w, cx, cy, n, sr)
            assert (mx, my, sad) == orc.motion_est(ref, cur, px, py, cx, cy, n, sr)
            blk = np.ascontiguousarray(cur[py:py + 8, px:px + 8])
            (mx, my), sad = me.motion_est(rbuf, stride, w, h, blk, 0, 8, cx, cy, 8, sr)
            assert (mx, my, sad) == orc.motion_est(ref, cur, px, py, cx, cy, 8, sr)
        # a 64x64 block: its last 4-byte unit is read as the aligned dword pair
        # that covers it, 4 bytes past the block in the drop-in's pinned buffer
        (mx, my), sad = me.motion_est(rbuf, stride, w, h, cur, 16 * w + 40, w, 37, 21, 64, 32)
        assert (mx, my, sad) == orc.motion_est(ref, cur, 40, 16, 37, 21, 64, 32)
        # same buffer, new content: next calls must see it
        rbuf[:, :w] = np.roll(ref, 5, 1)
        new = np.ascontiguousarray(rbuf[:, :w])
        (mx, my), sad = me.motion_est(rbuf, stride, w, h, cur, 0, w, 40, 40, 16, 24)
        assert (mx, my, sad) == orc.motion_est(new, cur, 0, 0, 40, 40, 16, 24)


def test_dropin_symbol_any_block_size(orc):
    """tc_motion_est through the drop-in symbol at block sizes the encoders
    never use (the reference accepts any), including sizes above 64 and a
    frame smaller than the block: every result equals the oracle's."""
    from tcodec_amd import me
    log("drop-in, any block size")
    rng = np.random.default_rng(33)
    ref, cur = moving_pair(120, 160, 4, shift=(3, -2))
    for n in (1, 3, 4, 5, 12, 24, 48, 65, 96, 120):
        for k in range(10):
            px = int(rng.integers(0, 160 - n + 1))
            py = int(rng.integers(0, 120 - n + 1))
            cx, cy = int(rng.integers(-10, 170)), int(rng.integers(-10, 130))
            sr = int(rng.choice([1, 8, 16, 24, 32, 64]))
            (mx, my), sad = me.motion_est(ref, 160, 160, 120, cur, py * 160 + px, 160, cx, cy, n, sr)
            assert (mx, my, sad) == orc.motion_est(ref, cur, px, py, cx, cy, n, sr), (n, k)
    # a 24x24 block against a 16x20 reference (no candidate fits)
    small = np.ascontiguousarray(ref[:20, :16])
    (mx, my), sad = me.motion_est(small, 16, 16, 20, cur, 0, 160, 3, 2, 24, 16)
    assert (mx, my, sad) == orc.motion_est(small, cur, 0, 0, 3, 2, 24, 16)


TILINGS = [(1, 1, 1, 2, 0, 85, 0), (1, 1, 2, 4, 4, 85, 0), (1, 1, 4, 8, 8, 85, 0),
           (1, 1, 1, 4, 8, 1, 0), (1, 1, 1, 4, 8, 37, 0), (1, 1, 1, 4, 4, 96, 0),
           (1, 1, 1, 4, 4, 85, 1), (1, 1, 1, 4, 4, 7, 1), (1, 1, 1, 4, 8, 85, 1),
           (1, 1, 1, 4, 8, 23, 1)]


@pytest.mark.parametrize("tiling", TILINGS, ids=lambda t: "w%d_%d_%d_%d_t%d_c%d_b%d" % t)
def test_gpu_tilings_agree_with_oracle(engine, orc, tiling):
    """Every launch shape (tile workgroups of 4 or 8 wavefronts and 1..96
    calls; or one workgroup per call with 1..8 wavefronts by block size) gives
    the oracle's answers and candidate counts, for CTU-major, reference-major
    and shuffled call orders."""
    from tcodec_amd import synth, workload
    log(f"tiling {tiling}")
    fr = synth.scene_video(256, 200, 5, seed=9, rects=6, pan=(0.75, 0.25))
    refs = [fr[i][0] for i in (0, 1, 3)]
    cur = fr[4][0]
    rps = [engine.pic(r) for r in refs]
    cp = engine.pic(cur)
    old = engine.tiling()
    engine.set_tiling(*tiling)
    try:
        for order in ("ctu", "ref_groups", "ref", "shuffled"):
            if order == "ctu":
                calls = workload.frame_calls_by_ctu(256, 200, rps, cp, workload.SR_B)
            elif order == "ref_groups":
                calls = workload.frame_calls_by_ctu(256, 200, rps, cp, workload.SR_B, ref_major=True)
            elif order == "ref":
                calls = workload.frame_calls(workload.qt_nodes(256, 200), rps, cp, workload.SR_P)
            else:
                calls = workload.frame_calls_by_ctu(256, 200, rps, cp, workload.SR_B)
                calls = calls[np.random.default_rng(3).permutation(len(calls))]
            res, st = engine.search(calls, stats=True)
            for k, rp in enumerate(rps):
                m = calls["ref_pic"] == rp
                want, wst = oracle_batch(orc, refs[k], cur, calls[m], stats=True)
                compare(res[m], want, calls[m], f"{tiling} {order} ref {k}")
                assert np.array_equal(st["int_evals"][m], wst[:, 0])
                assert np.array_equal(st["q6_evals"][m], wst[:, 2])
    finally:
        engine.set_tiling(**old.as_dict())
        for p in rps + [cp]:
            engine.free(p)


def test_gpu_autotune_keeps_results(engine, orc):
    """The autotuner only changes speed: after it runs, results are unchanged."""
    import torch
    from tcodec_amd import me, synth, workload
    log("autotune")
    fr = synth.scene_video(320, 240, 3, seed=2, rects=6)
    ref, cur = fr[0][0], fr[2][0]
    rp, cp = engine.pic(ref), engine.pic(cur)
    calls = workload.frame_calls_by_ctu(320, 240, [rp], cp, workload.SR_B)
    d_calls = torch.from_numpy(calls.view(np.int32).reshape(-1, 8).copy()).cuda()
    d_res = torch.zeros((len(calls), 4), dtype=torch.int32, device="cuda")
    old = engine.tiling()
    try:
        best, ms = engine.autotune(d_calls.data_ptr(), len(calls), d_res.data_ptr(), reps=1,
                                   stream=torch.cuda.current_stream().cuda_stream)
        assert ms > 0 and best.waves32 in (1, 2, 4) and best.waves64 in (2, 4, 8)
        assert best.tile_waves in (0, 4, 8)
        assert engine.tiling().as_dict() == best.as_dict()
        torch.cuda.synchronize()
        got = d_res.cpu().numpy().view(me.RESULT_DTYPE).ravel()
        compare(got, oracle_batch(orc, ref, cur, calls), calls, "autotune")
        compare(engine.search(calls), oracle_batch(orc, ref, cur, calls), calls, "after tune")
    finally:
        engine.set_tiling(**old.as_dict())
        engine.free(rp)
        engine.free(cp)


def test_gpu_kernel_timing(engine):
    """Per-kernel timing: one duration per launched kernel, in launch order."""
    from tcodec_amd import synth, workload
    fr = synth.scene_video(256, 192, 2, seed=4, rects=4)
    rp, cp = engine.pic(fr[0][0]), engine.pic(fr[1][0])
    calls = workload.frame_calls_by_ctu(256, 192, [rp], cp, workload.SR_P)
    old = engine.tiling()
    try:
        engine.set_timing(True)
        for tiling, count in (((1, 1, 1, 4, 8, 85, 0), 3), ((1, 1, 1, 4, 4, 85, 1), 1),
                              ((1, 1, 1, 4, 0, 85, 0), 4)):
            engine.set_tiling(*tiling)
            engine.search(calls)
            ms = engine.last_timing()
            assert len(ms) == count and all(t > 0 for t in ms), (tiling, ms)
    finally:
        engine.set_timing(False)
        engine.set_tiling(**old.as_dict())
        engine.free(rp)
        engine.free(cp)


def test_gpu_graph_replay(engine):
    """tcme_graph_*: the captured search replayed several times, the call list
    rewritten between replays and eager searches with synchronisations in
    between; every replay equals the eager result of the list it read, for the
    split launch (fork/join over side streams) and for tile_big."""
    import torch
    from tcodec_amd import me, synth, workload
    fr = synth.scene_video(256, 192, 4, seed=9, rects=4)
    pics = [engine.pic(f[0]) for f in fr]
    lists = [workload.frame_calls_by_ctu(256, 192, [pics[0], pics[1]], pics[2], workload.SR_B),
             workload.frame_calls_by_ctu(256, 192, [pics[1], pics[3]], pics[2], workload.SR_B),
             workload.frame_calls_by_ctu(256, 192, [pics[3], pics[0]], pics[1], workload.SR_B)]
    n = len(lists[0])
    assert all(len(x) == n for x in lists)
    want = [engine.search(x) for x in lists]
    host = [torch.from_numpy(x.view(np.int32).reshape(-1, 8).copy()) for x in lists]
    calls = host[0].cuda()
    res = torch.zeros((n, 4), dtype=torch.int32, device="cuda")
    stream = torch.cuda.current_stream()
    old = engine.tiling()
    try:
        for tiling in ((1, 1, 1, 4, 8, 85, 0), (1, 1, 1, 4, 4, 85, 1)):
            log(f"graph replay tiling {tiling}")
            engine.set_tiling(*tiling)
            g = engine.graph(calls.data_ptr(), n, res.data_ptr())
            try:
                for rep, k in enumerate((0, 1, 2, 1, 0, 2)):
                    calls.copy_(host[k])
                    res.fill_(-7)
                    g.launch(stream.cuda_stream)
                    torch.cuda.synchronize()
                    got = res.cpu().numpy().view(me.RESULT_DTYPE).ravel()
                    assert (got == want[k]).all(), (tiling, rep, k)
                    # eager work between replays
                    assert (engine.search(lists[(k + 1) % 3]) == want[(k + 1) % 3]).all()
            finally:
                g.close()
    finally:
        engine.set_tiling(**old.as_dict())
        for p in pics:
            engine.free(p)
