"""The TCMT ME trace format (tcodec_amd/trace.py): round trips of the golden
traces recorded from the reference encoder, malformed files, and replay of a
converted trace through the oracle (CPU)."""
import io

import numpy as np
import pytest

from helpers import golden_traces
from tcodec_amd import trace

CUR_BASE = 1000  # trace plane ids of current planes (reference planes keep their POC)


def to_tcmt(refs, curs, calls):
    """A golden fixture as a TCMT trace: every plane, then one CALLS record
    (golden columns: cur_id px py ref_poc cx cy blk sr w h mv_x mv_y sad)."""
    buf = io.BytesIO()
    w = trace.TraceWriter(buf)
    for poc, p in refs.items():
        w.plane(poc, p, trace.KIND_REF)
    for cid, p in curs.items():
        w.plane(CUR_BASE + cid, p, trace.KIND_CUR)
    rows = np.zeros((calls.shape[0], 12), np.int32)
    rows[:, 0] = CUR_BASE + calls[:, 0]
    rows[:, 1:3] = calls[:, 1:3]
    rows[:, 3] = calls[:, 3]
    rows[:, 4:8] = calls[:, 4:8]
    rows[:, 8:11] = calls[:, 10:13]
    w.calls(rows)
    return buf.getvalue(), rows


@pytest.mark.parametrize("fixture", golden_traces(), ids=lambda f: f[0])
def test_golden_round_trip(fixture):
    name, refs, curs, calls = fixture
    data, rows = to_tcmt(refs, curs, calls)
    recs = list(trace.read(io.BytesIO(data)))
    planes = [r for r in recs if isinstance(r, trace.Plane)]
    got = [r for r in recs if isinstance(r, trace.Calls)]
    assert len(planes) == len(refs) + len(curs) and len(got) == 1
    for p in planes:
        want = refs[p.id] if p.kind == trace.KIND_REF else curs[p.id - CUR_BASE]
        assert np.array_equal(p.data, want)
    assert np.array_equal(got[0].rows, rows)
    # header (16) + per plane 8 + 16 + w*h + calls 8 + 4 + 48 per call
    size = 16 + sum(24 + p.data.size for p in planes) + 12 + 48 * rows.shape[0]
    assert len(data) == size


def test_malformed_traces_raise():
    data, _ = to_tcmt({0: np.zeros((8, 8), np.uint8)}, {}, np.zeros((0, 16), np.int32))
    with pytest.raises(ValueError):
        list(trace.read(io.BytesIO(b"XXXX" + data[4:])))
    with pytest.raises(ValueError):
        list(trace.read(io.BytesIO(data[:4] + b"\x02\x00\x00\x00" + data[8:])))
    with pytest.raises(ValueError):
        list(trace.read(io.BytesIO(data[:-5])))
    # a record of an unknown tag is skipped
    extra = data + (9).to_bytes(4, "little") + (3).to_bytes(4, "little") + b"abc"
    assert len(list(trace.read(io.BytesIO(extra)))) == 2


@pytest.mark.parametrize("fixture", golden_traces()[:1], ids=lambda f: f[0])
def test_converted_trace_replays_on_the_oracle(orc, fixture):
    """Replaying the converted trace's searches through the oracle gives the
    recorded (reference) results: the record semantics are the call's."""
    name, refs, curs, calls = fixture
    data, _ = to_tcmt(refs, curs, calls)
    planes = {}
    n = 0
    for rec in trace.read(io.BytesIO(data)):
        if isinstance(rec, trace.Plane):
            planes[rec.id] = rec.data
            continue
        for r in rec.rows[:600]:
            got = orc.motion_est(planes[int(r[3])], planes[int(r[0])], int(r[1]), int(r[2]),
                                 int(r[4]), int(r[5]), int(r[6]), int(r[7]))
            assert tuple(got) == tuple(int(v) for v in r[8:11]), r
            n += 1
    assert n > 100
