"""ME call traces: a stable on-disk record of searches, for replay and parity.

A trace holds the planes a sequence of motion searches read and the searches
themselves with their results, so that the same searches can be run again on
another host or build (``replay``) and compared call for call.  The serving
loop records one with ``Server(trace=TraceWriter(path))``; the reference
encoder's calls can be converted from the golden fixtures.

Format TCMT, version 1 (all integers little-endian int32/uint32)::

    header   b"TCMT", u32 version = 1, u32 flags = 0, u32 reserved = 0
    record   u32 tag, u32 payload bytes, payload
      tag 1  PLANE: i32 id, i32 kind, i32 w, i32 h, then w*h bytes, row-major
             (kind 0: a reference plane, searched in; 1: a current plane)
      tag 2  CALLS: i32 count, then count rows of 12 i32:
             cur_id px py ref_id cx cy blk sr mv_x mv_y sad status

A PLANE record defines plane ``id`` (or replaces its content: ids are slots,
like the encoder's picture buffer); a CALLS record refers to the planes as
last defined before it.  ``cur_id``/``ref_id``, ``px``/``py`` (block origin),
``cx``/``cy`` (full-pel centre), ``blk`` and ``sr`` are the arguments of the
reference's ``tc_motion_est`` (reference src/motion.c:384-389) with the two
planes named by id; ``mv_x``/``mv_y`` (absolute quarter-pel) and ``sad`` are its
results and ``status`` the batch API's per-call status (0 for a valid call).
Unknown tags are skipped by readers, so later versions may add records.
"""
from __future__ import annotations

import struct
from dataclasses import dataclass
from typing import BinaryIO, Iterator

import numpy as np

MAGIC = b"TCMT"
VERSION = 1
TAG_PLANE, TAG_CALLS = 1, 2
CALL_FIELDS = ("cur_id", "px", "py", "ref_id", "cx", "cy", "blk", "sr", "mv_x", "mv_y", "sad",
               "status")
KIND_REF, KIND_CUR = 0, 1


@dataclass
class Plane:
    id: int
    kind: int
    data: np.ndarray   # uint8 [h, w]


@dataclass
class Calls:
    rows: np.ndarray   # int32 [count, 12], columns CALL_FIELDS


class TraceWriter:
    def __init__(self, path_or_file):
        self.f: BinaryIO = open(path_or_file, "wb") if isinstance(path_or_file, str) \
            else path_or_file
        self.f.write(MAGIC + struct.pack("<III", VERSION, 0, 0))
        self.ncalls = 0

    def plane(self, pid: int, plane: np.ndarray, kind: int = KIND_REF) -> None:
        plane = np.ascontiguousarray(plane, np.uint8)
        h, w = plane.shape
        self.f.write(struct.pack("<II", TAG_PLANE, 16 + w * h))
        self.f.write(struct.pack("<iiii", pid, kind, w, h))
        self.f.write(plane.tobytes())

    def calls(self, rows: np.ndarray) -> None:
        rows = np.ascontiguousarray(rows, "<i4").reshape(-1, len(CALL_FIELDS))
        self.f.write(struct.pack("<II", TAG_CALLS, 4 + rows.nbytes))
        self.f.write(struct.pack("<i", rows.shape[0]))
        self.f.write(rows.tobytes())
        self.ncalls += rows.shape[0]

    def close(self) -> None:
        self.f.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def rows_from(calls: np.ndarray, results: np.ndarray, cur_ids=None, ref_ids=None) -> np.ndarray:
    """CALLS rows from a batch-API call list and its results: the picture ids
    in the call list are replaced by trace plane ids (``cur_ids``/``ref_ids``
    map picture id -> plane id; None keeps the picture ids)."""
    calls = np.ascontiguousarray(calls).view(np.int32).reshape(-1, 8)
    res = np.ascontiguousarray(results).view(np.int32).reshape(-1, 4)
    out = np.empty((calls.shape[0], len(CALL_FIELDS)), np.int32)
    out[:, 0] = calls[:, 1] if cur_ids is None else [cur_ids[int(p)] for p in calls[:, 1]]
    out[:, 1:3] = calls[:, 2:4]
    out[:, 3] = calls[:, 0] if ref_ids is None else [ref_ids[int(p)] for p in calls[:, 0]]
    out[:, 4:8] = calls[:, 4:8]
    out[:, 8:12] = res
    return out


def read(path_or_file) -> Iterator[Plane | Calls]:
    """The records of a trace, in order.  Raises ValueError on a bad header
    or a truncated record."""
    f = open(path_or_file, "rb") if isinstance(path_or_file, str) else path_or_file
    try:
        head = f.read(16)
        if len(head) != 16 or head[:4] != MAGIC:
            raise ValueError("not a TCMT trace")
        version = struct.unpack("<I", head[4:8])[0]
        if version != VERSION:
            raise ValueError(f"TCMT version {version} (this reader knows {VERSION})")
        while True:
            rh = f.read(8)
            if not rh:
                return
            if len(rh) != 8:
                raise ValueError("truncated record header")
            tag, n = struct.unpack("<II", rh)
            body = f.read(n)
            if len(body) != n:
                raise ValueError("truncated record")
            if tag == TAG_PLANE:
                pid, kind, w, h = struct.unpack("<iiii", body[:16])
                if w <= 0 or h <= 0 or n != 16 + w * h:
                    raise ValueError(f"bad plane record {pid}: {w}x{h}, {n} bytes")
                yield Plane(pid, kind, np.frombuffer(body[16:], np.uint8).reshape(h, w).copy())
            elif tag == TAG_CALLS:
                count = struct.unpack("<i", body[:4])[0]
                if count < 0 or n != 4 + count * 4 * len(CALL_FIELDS):
                    raise ValueError(f"bad calls record: {count} calls, {n} bytes")
                yield Calls(np.frombuffer(body[4:], "<i4").reshape(count, len(CALL_FIELDS))
                            .astype(np.int32))
    finally:
        if isinstance(path_or_file, str):
            f.close()


def replay(eng, path_or_file, batch: int = 1 << 20) -> dict:
    """Runs every search of a trace on the GPU (``me.Engine``) and compares
    the results with the recorded ones.  Returns counts and the first
    mismatching row (recorded, then got)."""
    from . import me
    slots: dict[int, int] = {}
    sizes: dict[int, tuple[int, int]] = {}
    ncalls = nbad = 0
    first = None
    try:
        for rec in read(path_or_file):
            if isinstance(rec, Plane):
                h, w = rec.data.shape
                if rec.id in slots and sizes[rec.id] == (w, h):
                    eng.upload(slots[rec.id], rec.data)
                else:
                    if rec.id in slots:
                        eng.free(slots.pop(rec.id))
                    slots[rec.id] = eng.pic(rec.data)
                    sizes[rec.id] = (w, h)
                continue
            rows = rec.rows
            for s in range(0, rows.shape[0], batch):
                r = rows[s:s + batch]
                calls = me.make_calls([slots[int(i)] for i in r[:, 3]],
                                      [slots[int(i)] for i in r[:, 0]], r[:, 1], r[:, 2],
                                      r[:, 4], r[:, 5], r[:, 6], r[:, 7])
                got = eng.search(calls).view(np.int32).reshape(-1, 4)
                bad = np.nonzero((got != r[:, 8:12]).any(1))[0]
                if bad.size and first is None:
                    first = (r[bad[0]].tolist(), got[bad[0]].tolist())
                nbad += int(bad.size)
                ncalls += r.shape[0]
    finally:
        for p in slots.values():
            eng.free(p)
    return {"calls": ncalls, "mismatches": nbad, "first_mismatch": first}
