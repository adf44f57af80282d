"""The served and batched forms: every stream's results are the same alone
and in the mix, and match the oracle."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_served_results_same_alone_and_in_mix(engine, orc):
    from tcodec_amd import serve, workload
    print("serve mix", flush=True)
    clips = {"a": serve.Clip(160, 128, 5, seed=11), "b": serve.Clip(256, 144, 5, seed=12),
             "c": serve.Clip(320, 240, 5, seed=13, kind="gradient")}
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
