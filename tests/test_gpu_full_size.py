"""Full-size frames: every call of a 1080p frame (the bench workload: 171,600
calls against 4 references) and of a 2160p frame (2 references), searched by
the default one-launch tile kernel and by the per-call kernels (a separate
implementation of the walk, `tile_waves = 0`), must agree call for call,
candidate counts included; a seeded sample of calls is checked against the
oracle (which would take minutes for the whole frame)."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def check_frame(engine, orc, planes, ref_pocs, cur_poc, sr, sample, seed):
    from tcodec_amd import me, workload
    h, w = planes[cur_poc].shape
    pics = {p: engine.pic(planes[p]) for p in list(ref_pocs) + [cur_poc]}
    calls = workload.frame_calls_by_ctu(w, h, [pics[p] for p in ref_pocs], pics[cur_poc], sr,
                                        ref_major=True)
    old = engine.tiling()
    try:
        engine.set_tiling(1, 1, 1, 4, 8, 85, 1)
        res_t, st_t = engine.search(calls, stats=True)
        engine.set_tiling(1, 1, 1, 4, 0, 85, 0)
        res_p, st_p = engine.search(calls, stats=True)
    finally:
        engine.set_tiling(**old.as_dict())
        for p in pics.values():
            engine.free(p)
    assert (res_t["status"] == 0).all()
    diff = np.nonzero(res_t.view(np.int32).reshape(-1, 4) != res_p.view(np.int32).reshape(-1, 4))[0]
    assert diff.size == 0, f"{w}x{h}: tile and per-call kernels differ on {np.unique(diff).size} calls"
    assert np.array_equal(st_t.view(np.int32), st_p.view(np.int32))
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(len(calls), sample, replace=False))
    for r in ref_pocs:
        m = idx[calls["ref_pic"][idx] == pics[r]]
        arr = np.stack([calls["px"][m], calls["py"][m], calls["cx"][m], calls["cy"][m],
                        calls["blk"][m], calls["sr"][m]], 1)
        want = orc.motion_est_batch(planes[r], planes[cur_poc], arr)
        got = np.stack([res_t["mv_x"][m], res_t["mv_y"][m], res_t["sad"][m]], 1)
        assert np.array_equal(got, want), (w, h, r)
    return len(calls)


def test_gpu_1080p_bench_frame(engine, orc):
    from tcodec_amd import synth, workload
    log("1080p frame of the bench workload")
    order = workload.gop4_coding_order(13)
    poc, kind, dpb = order[8]
    assert len(dpb) == 4
    scene = synth.Scene(1920, 1080, seed=2026, rects=48, pan=(0.75, 0.25))
    planes = {p: scene.frame(p)[0] for p in list(dpb) + [poc]}
    n = check_frame(engine, orc, planes, dpb, poc,
                    workload.SR_B if kind == "B" else workload.SR_P, 3000, 1)
    assert n == 171600


def test_gpu_2160p_frame(engine, orc):
    from tcodec_amd import synth, workload
    log("2160p frame, 2 references")
    frames = synth.gradient_video(3840, 2160, 3, seed=7)
    planes = {t: frames[t][0] for t in range(3)}
    n = check_frame(engine, orc, planes, [0, 1], 2, workload.SR_P, 2000, 2)
    assert n == 2 * len(workload.qt_nodes(3840, 2160))
