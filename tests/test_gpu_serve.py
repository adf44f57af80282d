"""The served and batched forms: every stream's results are the same alone
and in the mix, and match the oracle."""
import sys
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def oracle_check(orc, stream, pos, got, sample=None, seed=0):
    """Results of coded picture `pos` of a served stream (every call of its
    frame template, in order) against the oracle; `sample` calls drawn with a
    seeded generator, or all of them.  The oracle runs on a thread pool (one
    foreign call per chunk, so the threads run in parallel)."""
    from tcodec_amd import serve, workload
    poc, kind, dpb = stream.order[pos]
    calls = serve.frame_template(stream.w, stream.h, len(dpb),
                                 workload.SR_B if kind == "B" else workload.SR_P)
    assert got.shape[0] == len(calls)
    clip = stream.clip

    def plane(p):
        t = (stream.offset + p) % clip.frames
        return clip.host[t] if clip.host is not None else clip.dev[t].cpu().numpy()

    cur = plane(poc)
    arr = np.stack([calls["px"], calls["py"], calls["cx"], calls["cy"], calls["blk"],
                    calls["sr"]], 1).astype(np.int32)
    idx = np.arange(len(calls))
    if sample is not None and sample < len(calls):
        idx = np.sort(np.random.default_rng(seed).choice(len(calls), sample, replace=False))
    jobs = []
    for k, p in enumerate(dpb):
        ref = plane(p)
        sel = idx[calls["ref_pic"][idx] == k]
        for c0 in range(0, sel.size, 4096):
            jobs.append((ref, sel[c0:c0 + 4096]))

    def run(job):
        ref, sel = job
        want = orc.motion_est_batch(ref, cur, arr[sel])
        bad = np.nonzero((got[sel, :3] != want).any(1) | (got[sel, 3] != 0))[0]
        return None if bad.size == 0 else (int(sel[bad[0]]), got[sel[bad[0]]], want[bad[0]])

    with ThreadPoolExecutor(16) as pool:
        for r in pool.map(run, jobs):
            assert r is None, (stream.sid, stream.w, stream.h, poc, "call", r)
    return idx.size


def test_served_results_same_alone_and_in_mix(engine, orc):
    from tcodec_amd import serve, workload
    print("serve mix", flush=True)
    # 12-picture clips: a stream (at most 9 pictures, offset up to 4) never
    # meets the same source picture as current and reference
    clips = {"a": serve.Clip(160, 128, 12, seed=11), "b": serve.Clip(256, 144, 12, seed=12),
             "c": serve.Clip(320, 240, 12, seed=13, kind="gradient")}
    mix = [serve.Stream(0, 160, 128, 9, "B", clips["a"], offset=0, arrival=0.0),
           serve.Stream(1, 256, 144, 7, "P", clips["b"], offset=2, arrival=0.0),
           serve.Stream(2, 320, 240, 9, "B", clips["c"], offset=1, arrival=0.02),
           serve.Stream(3, 160, 128, 5, "P", clips["a"], offset=3, arrival=0.03),
           serve.Stream(4, 256, 144, 9, "B", clips["b"], offset=4, arrival=0.05)]
    # a budget that admits three streams at once, so later ones wait for slots
    budget = 3 * serve.stream_bytes(320, 240)
    srv = serve.Server(engine, hbm_budget=budget, keep_results=True)
    for s in mix:
        srv.submit(s)
    srv.run()
    assert len(srv.finished) == len(mix) and srv.used == 0
    for s in srv.finished:
        alone = serve.run_alone(engine, s)
        assert len(alone) == len(s.results)
        for a, b in zip(alone, s.results):
            assert np.array_equal(a, b), s.sid
        assert (np.concatenate(s.results)[:, 3] == 0).all()
    # one stream's last picture against the oracle, call for call
    s = next(x for x in srv.finished if x.sid == 0)
    poc, kind, dpb = s.order[-1]
    calls = serve.frame_template(s.w, s.h, len(dpb),
                                 workload.SR_B if kind == "B" else workload.SR_P)
    plane = lambda p: s.clip.host[(s.offset + p) % s.clip.frames]
    got = s.results[-1]
    for k, p in enumerate(dpb):
        m = calls["ref_pic"] == k
        arr = np.stack([calls["px"], calls["py"], calls["cx"], calls["cy"], calls["blk"],
                        calls["sr"]], 1)[m]
        want = orc.motion_est_batch(plane(p), plane(poc), arr)
        assert np.array_equal(got[m][:, :3], want), (poc, p)


def test_batched_one_launch_equals_per_picture(engine):
    from tcodec_amd import serve
    print("batched vs per picture", flush=True)
    clips = [serve.DeviceGradientClip(w, h, 7, seed=s)
             for s, (w, h) in enumerate([(320, 240), (640, 360), (320, 240), (1280, 720)])]

    def run(per_stream):
        srv = serve.Server(engine, hbm_budget=1 << 40, max_tick_calls=1 << 40,
                           keep_results=True, per_stream_launch=per_stream)
        for i, c in enumerate(clips):
            srv.submit(serve.Stream(i, c.w, c.h, 7, "P", c))
        srv.run()
        return {s.sid: s.results for s in srv.finished}

    a, b = run(False), run(True)
    for sid in a:
        assert len(a[sid]) == len(b[sid]) == 6
        for x, y in zip(a[sid], b[sid]):
            assert np.array_equal(x, y), sid


def test_gpu_trace_replays(engine, tmp_path):
    """TCMT traces replayed on the GPU: the golden traces of the reference
    encoder, and a serving session recorded by the server itself (streams
    admitted and released, so picture ids are reused as trace slots)."""
    import io

    from helpers import golden_traces
    from tcodec_amd import serve, trace
    from test_trace import to_tcmt
    print("trace replay: golden", flush=True)
    for name, refs, curs, calls in golden_traces():
        data, _ = to_tcmt(refs, curs, calls)
        r = trace.replay(engine, io.BytesIO(data))
        assert r["calls"] == calls.shape[0] and r["mismatches"] == 0, (name, r)
    print("trace replay: served session", flush=True)
    clip_a = serve.Clip(160, 128, 5, seed=21)
    clip_b = serve.DeviceGradientClip(256, 144, 6, seed=22)
    path = str(tmp_path / "served.tcmt")
    with trace.TraceWriter(path) as tw:
        srv = serve.Server(engine, hbm_budget=2 * serve.stream_bytes(256, 144), trace=tw)
        srv.submit(serve.Stream(0, 160, 128, 5, "B", clip_a, offset=1, arrival=0.0))
        srv.submit(serve.Stream(1, 256, 144, 6, "P", clip_b, offset=0, arrival=0.0))
        srv.submit(serve.Stream(2, 160, 128, 5, "P", clip_a, offset=2, arrival=0.01))
        srv.run()
    r = trace.replay(engine, path)
    assert r["calls"] == srv.calls and r["calls"] > 0 and r["mismatches"] == 0, r


MIX4 = [("240p", 16), ("720p", 24), ("1080p", 16), ("2160p", 8)]


def test_config4_64_stream_mix_one_launch_per_time_step(engine, orc):
    """The batched configuration: 64 P-only streams of uneven size (16 x
    320x240, 24 x 1280x720, 16 x 1920x1080, 8 x 3840x2160), gradient source
    with a seed per stream, 4 references.  Every time step is ONE tcme_search
    over every stream's next picture.  The last step (each picture with a
    full 4-slot buffer, 10.2 M calls) is checked: every call of the 240p and
    720p streams against the oracle, a seeded sample of every 1080p and 2160p
    stream; and every picture of every stream is equal (per-picture digests)
    to the same streams searched with the per-call launch shape and with one
    launch per picture."""
    from tcodec_amd import serve
    nf = 6                                 # intra + 4 P to fill the buffer + 1 checked
    clips, sid = [], 0
    for name, count in MIX4:
        w, h = serve.RES[name]
        for _ in range(count):
            clips.append((sid, w, h, serve.DeviceGradientClip(w, h, nf, seed=1000 + sid)))
            sid += 1

    def run(per_stream=False):
        srv = serve.Server(engine, hbm_budget=1 << 62, max_tick_calls=1 << 62,
                           keep_results="digest", per_stream_launch=per_stream)
        for sid, w, h, clip in clips:
            srv.submit(serve.Stream(sid, w, h, nf, "P", clip, keep_at={nf - 1}))
        launches = []
        orig = engine.tick_search

        def counted(recs, res_ptr, res_cap, stream=0):
            launches.append(res_cap)          # one tcme_tick_search: one search launch
            return orig(recs, res_ptr, res_cap, stream)

        engine.tick_search = counted
        try:
            srv.run()
        finally:
            del engine.tick_search
        assert len(srv.finished) == 64 and srv.ticks == nf and srv.max_active == 64
        return srv, launches

    log("config 4: batched, tile shape")
    srv, launches = run()
    assert len(launches) == nf - 1         # one launch per time step (intra step: none)
    assert launches[-1] > 10_000_000, launches
    t0 = engine.tiling()
    log("config 4: per-call launch shape")
    engine.set_tiling(1, 1, 1, 4, 0, 85, 0)
    try:
        srv_pc, _ = run()
    finally:
        engine.set_tiling(t0.waves8, t0.waves16, t0.waves32, t0.waves64, t0.tile_waves,
                          t0.tile_chunk, t0.tile_big)
    log("config 4: one launch per picture")
    srv_pp, launches_pp = run(per_stream=True)
    assert len(launches_pp) == 64 * (nf - 1)
    a = {s.sid: s for s in srv.finished}
    for other in (srv_pc, srv_pp):
        for s in other.finished:
            assert s.digests == a[s.sid].digests, s.sid
            assert np.array_equal(s.kept[nf - 1], a[s.sid].kept[nf - 1]), s.sid
    checked = 0
    for s in srv.finished:
        full = s.w <= 1280
        log(f"config 4: oracle stream {s.sid} {s.w}x{s.h} {'all' if full else 'sample'}")
        checked += oracle_check(orc, s, nf - 1, s.kept[nf - 1], None if full else 3000,
                                seed=s.sid)
    log(f"config 4: {checked} calls checked against the oracle")


def test_config5_served_session_mix(engine, orc):
    """The serving configuration at test scale: Poisson arrivals at 40
    streams/s for 1 s, resolutions drawn 30/40/25/5 % from 320x240 / 1280x720
    / 1920x1080 / 3840x2160, 30..40 pictures each, B-GOP4, 4 references,
    sources from the scene generator (device-rendered clips, every picture
    distinct), under an HBM budget that makes later requests wait for slots.
    Every picture's results (digest) are the same served alone; one picture
    per stream is checked against the oracle on a seeded sample."""
    from tcodec_amd import serve
    clips = {name: serve.DeviceSceneClip(w, h, 48, seed=77 + i)
             for i, (name, (w, h)) in enumerate(serve.RES.items())}
    reqs = serve.poisson_requests(40.0, 1.0, seed=11, clips=clips, frames=(30, 40))
    sizes = {(r.w, r.h) for r in reqs}
    assert sizes == set(serve.RES.values()), sizes
    for r in reqs:
        r.keep_at = {len(r.order) - 1}     # the last coded picture (a B picture)
    budget = 3 * serve.stream_bytes(3840, 2160)
    srv = serve.Server(engine, hbm_budget=budget, keep_results="digest")
    for r in reqs:
        srv.submit(r)
    log(f"config 5: {len(reqs)} requests")
    srv.run()
    assert len(srv.finished) == len(reqs) and not srv.dropped and srv.used == 0
    assert srv.max_active > 1 and srv.admitted_while_busy > 0
    log(f"config 5: {srv.frames} pictures in {srv.ticks} launches, "
        f"at most {srv.max_active} streams at once, {srv.admitted_while_busy} late admissions")
    for s in srv.finished:
        dg, kept = serve.run_alone(engine, s, keep="digest")
        assert dg == s.digests, s.sid
        pos = len(s.order) - 1
        assert np.array_equal(kept[pos], s.kept[pos]), s.sid
        assert s.order[pos][1] == "B"
        oracle_check(orc, s, pos, s.kept[pos], 1500, seed=s.sid)
    log("config 5: alone == mix for every stream; oracle samples equal")
