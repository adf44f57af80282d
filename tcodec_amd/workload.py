"""Motion-estimation call sets for throughput and latency runs.

The reference encoder issues ``tc_motion_est`` from its RDO (reference
src/encoder.c:807, 834, 928, 982, 1004, 1035) at every quadtree node it
evaluates, centred on a predictor that depends on earlier RDO decisions.
Throughput runs cannot wait for a CPU RDO, so they use a fixed, documented
call set per frame that the GPU can search without it:

* every in-frame quadtree node of sizes 64, 32, 16 and 8 (a node is in frame
  when it lies wholly inside the picture, reference src/encoder.c:1297-1306),
* against every picture in the frame's 4-slot decoded picture buffer
  (reference src/encoder.c:3085-3097; B-frames of the GOP4 ladder,
  src/encoder.c:3123-3139, hold their forward and backward anchors there),
* centred on the block's own position (zero predictor),
* with the medium-preset search ranges: 24 for P, 32 for B
  (reference src/encoder.c:801, 926).
"""
from __future__ import annotations

import numpy as np

from .me import CALL_DTYPE

SIZES = (64, 32, 16, 8)
SR_P, SR_B = 24, 32


def qt_nodes(w: int, h: int, sizes=SIZES) -> np.ndarray:
    """int32 [N,3] = (px, py, n) for every in-frame node, largest first."""
    out = []
    for n in sizes:
        ys, xs = np.mgrid[0:h - n + 1:n, 0:w - n + 1:n]
        out.append(np.stack([xs.ravel(), ys.ravel(), np.full(xs.size, n)], 1))
    return np.concatenate(out).astype(np.int32)


def bf_schedule_poc(emit_pos: int) -> int:
    """Display POC of the emit_pos-th coded picture in the GOP4 ladder
    (reference src/encoder.c:3133-3140): A0, then A4k, B4k-2, B4k-3, B4k-1."""
    if emit_pos == 0:
        return 0
    k = (emit_pos + 3) // 4
    c = (emit_pos - 1) % 4
    return 4 * k + (0, -2, -3, -1)[c]


def gop4_coding_order(frames: int):
    """[(poc, kind, dpb_pocs)] in coding order for `frames` display frames
    (a multiple of 4 plus one); kind is 'I', 'P' or 'B'; dpb_pocs is the
    4-slot buffer, most recent first, at the time the picture is coded."""
    assert frames >= 1 and (frames - 1) % 4 == 0, "GOP4 run needs 4k+1 frames"
    dpb: list[int] = []
    order = []
    for e in range(frames):
        poc = bf_schedule_poc(e)
        kind = "I" if poc == 0 else ("P" if poc % 4 == 0 else "B")
        order.append((poc, kind, list(dpb)))
        dpb = ([poc] + dpb)[:4]
    return order


def p_coding_order(frames: int):
    """P-only: [(poc, kind, dpb_pocs)] with the same 4-slot buffer."""
    dpb: list[int] = []
    order = []
    for poc in range(frames):
        order.append((poc, "I" if poc == 0 else "P", list(dpb)))
        dpb = ([poc] + dpb)[:4]
    return order


def frame_calls(nodes: np.ndarray, ref_pics, cur_pic: int, sr: int) -> np.ndarray:
    """Calls for one frame: every node against every reference picture."""
    nodes = np.asarray(nodes, np.int32)
    refs = np.asarray(list(ref_pics), np.int32)
    m = nodes.shape[0]
    out = np.zeros(m * refs.size, CALL_DTYPE)
    out["ref_pic"] = np.repeat(refs, m)
    out["cur_pic"] = cur_pic
    out["px"] = np.tile(nodes[:, 0], refs.size)
    out["py"] = np.tile(nodes[:, 1], refs.size)
    out["cx"] = out["px"]
    out["cy"] = out["py"]
    out["blk"] = np.tile(nodes[:, 2], refs.size)
    out["sr"] = sr
    return out


def algo_bytes(stats: np.ndarray, blk: np.ndarray) -> int:
    """Algorithmic bytes of searched calls: an in-frame integer SAD reads
    2*n^2 samples, a 6-tap quarter-pel SAD n^2 + (n+5)^2, a bilinear one
    n^2 + (n+1)^2; rejected candidates read nothing."""
    n = np.asarray(blk, np.int64)
    nn = n * n
    return int((stats["int_evals"].astype(np.int64) * 2 * nn +
                stats["q6_evals"].astype(np.int64) * (nn + (n + 5) ** 2) +
                stats["qbil_evals"].astype(np.int64) * (nn + (n + 1) ** 2)).sum())


def frame_calls_by_ctu(w: int, h: int, ref_pics, cur_pic: int, sr: int,
                       sizes=SIZES, ref_major: bool = False) -> np.ndarray:
    """The same call set as ``frame_calls(qt_nodes(w, h), ...)`` ordered for
    the chunked kernel: grouped by (CTU, reference), complete CTUs (85 nodes)
    first and partial edge CTUs last, largest blocks first inside a group, so
    a chunk of 85 calls is exactly one CTU searched against one reference.
    Groups follow CTU order, a CTU's references in turn; with ``ref_major``
    every CTU against the first reference comes first, then the second, ...
    (neighbouring chunks then search neighbouring CTUs in the same picture)."""
    nodes = qt_nodes(w, h, sizes)
    ctu = (nodes[:, 1] // 64) * ((w + 63) // 64) + nodes[:, 0] // 64
    counts = np.bincount(ctu, minlength=int(ctu.max()) + 1)
    partial = (counts[ctu] != 85).astype(np.int64)
    refs = np.asarray(list(ref_pics), np.int32)
    m = nodes.shape[0]
    r_idx = np.repeat(np.arange(refs.size), m)
    nd = np.tile(np.arange(m), refs.size)
    if ref_major:
        order = np.lexsort((nodes[nd, 0], nodes[nd, 1], -nodes[nd, 2], ctu[nd], r_idx, partial[nd]))
    else:
        order = np.lexsort((nodes[nd, 0], nodes[nd, 1], -nodes[nd, 2], r_idx, ctu[nd], partial[nd]))
    out = np.zeros(m * refs.size, CALL_DTYPE)
    sel = nd[order]
    out["ref_pic"] = refs[r_idx[order]]
    out["cur_pic"] = cur_pic
    out["px"] = nodes[sel, 0]
    out["py"] = nodes[sel, 1]
    out["cx"] = nodes[sel, 0]
    out["cy"] = nodes[sel, 1]
    out["blk"] = nodes[sel, 2]
    out["sr"] = sr
    return out
