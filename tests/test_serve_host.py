"""Host logic of the serving loop (no GPU): admission by HBM budget."""
import time

from tcodec_amd import serve


def _stream(sid, w, h, arrival):
    return serve.Stream(sid, w, h, 5, "P", clip=None, arrival=arrival)


def test_request_larger_than_budget_is_refused_not_blocking():
    """A request whose pictures exceed the whole budget is refused (recorded
    in `dropped`) instead of blocking the queue forever, and run() returns."""
    srv = serve.Server(eng=None, hbm_budget=serve.stream_bytes(320, 240) - 1)
    srv.submit(_stream(0, 3840, 2160, 0.0))
    srv.submit(_stream(1, 1920, 1080, 0.01))
    t0 = time.perf_counter()
    srv.run()
    assert time.perf_counter() - t0 < 5
    assert [s.sid for s in srv.dropped] == [0, 1]
    assert not srv.active and not srv.queue and not srv.finished and srv.used == 0


def test_stream_bytes_counts_five_resident_pictures():
    assert serve.stream_bytes(1920, 1080) == 5 * serve.pic_bytes(1920, 1080)
    assert serve.pic_bytes(320, 240) == 320 * 240 * 9 + 512


def test_library_step_layout_equals_python_template():
    """tcme_tick_layout (the call list the library builds on the device for a
    time step) is exactly the Python frame template the oracle checks use:
    complete CTUs first, reference-major, largest blocks first."""
    import numpy as np

    from tcodec_amd import me
    for (w, h, refs, sr) in [(320, 240, 4, 24), (1280, 720, 2, 32), (200, 136, 1, 16),
                             (64, 64, 3, 24), (1920, 1080, 4, 32), (40, 24, 1, 24)]:
        a = me.tick_layout(w, h, refs, sr)
        b = serve.frame_template(w, h, refs, sr)
        assert len(a) == len(b), (w, h)
        for f in me.CALL_DTYPE.names:
            if f != "cur_pic":
                assert np.array_equal(a[f], b[f]), (w, h, f)
        assert (a["cur_pic"] == -1).all()
