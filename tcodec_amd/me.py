"""Host-side mirror of TCodec's motion-estimation interface over libtcme.so.

The reference exposes one C function for this path,
``tc_motion_est(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x,
center_y, blk_size, search_range, &best_sad)`` (reference
include/tcodec_common.h:454-459, src/motion.c:384-485).  ``motion_est`` below
calls the library's drop-in symbol with the same arguments and meaning.  The
batch class ``Engine`` drives the ``tcme_*`` entry points: device-resident
pictures and many calls per launch.

Nothing here computes motion estimation: every result comes from the gfx950
kernels.  Loading fails loudly when the library is missing or no GPU exists.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libtcme.so")
ROOT = os.path.dirname(HERE)

# Every symbol include/tcme.h declares.
EXPORTS = (
    "tc_motion_est", "tcme_last_error", "tcme_version", "tcme_create", "tcme_destroy",
    "tcme_pic_create", "tcme_pic_destroy", "tcme_pic_upload", "tcme_pic_upload_device",
    "tcme_pic_luma", "tcme_pic_finish", "tcme_search", "tcme_search_host",
    "tcme_set_tiling", "tcme_get_tiling", "tcme_autotune", "tcme_set_timing",
    "tcme_last_timing", "tcme_pic_create_plane", "tcme_predict_luma", "tcme_predict_chroma",
    "tcme_global_mv", "tcme_sad", "tcme_graph_create", "tcme_graph_launch",
    "tcme_graph_destroy", "tcme_prepare", "tcme_frame_create", "tcme_frame_destroy",
    "tcme_frame_begin", "tcme_frame_end", "tcme_frame_motion_est", "tcme_frame_get_stats",
    "tcme_global_mv_static", "tcme_search_chain", "tcme_tick_search", "tcme_tick_layout",
)

CALL_DTYPE = np.dtype([("ref_pic", "<i4"), ("cur_pic", "<i4"), ("px", "<i4"), ("py", "<i4"),
                       ("cx", "<i4"), ("cy", "<i4"), ("blk", "<i4"), ("sr", "<i4")])
RESULT_DTYPE = np.dtype([("mv_x", "<i4"), ("mv_y", "<i4"), ("sad", "<i4"), ("status", "<i4")])
PRED_DTYPE = np.dtype([("pic", "<i4"), ("mv_x", "<i4"), ("mv_y", "<i4"), ("blk", "<i4"),
                       ("out_off", "<i8"), ("out_stride", "<i4"), ("reserved", "<i4")])
SAD_DTYPE = np.dtype([("pic", "<i4"), ("x", "<i4"), ("y", "<i4"), ("w", "<i4"), ("h", "<i4"),
                      ("other_pic", "<i4"), ("ox", "<i4"), ("oy", "<i4"), ("buf_off", "<i8"),
                      ("buf_stride", "<i4"), ("reserved", "<i4")])
CHAIN_FRAME_DTYPE = np.dtype([("cur_pic", "<i4"), ("nrefs", "<i4"), ("ref_pic", "<i4", (4,)),
                              ("sr", "<i4"), ("lambda", "<i4")])
CHAIN_CALL_DTYPE = np.dtype([(k, "<i4") for k in ("ref", "px", "py", "cx", "cy", "blk", "mv_x",
                                                   "mv_y", "sad", "reserved")])
CHAIN_MAX_CALLS = 676
TICK_PIC_DTYPE = np.dtype([("cur_pic", "<i4"), ("nrefs", "<i4"), ("sr", "<i4"),
                           ("ref_pic", "<i4", (4,))])
STATS_DTYPE = np.dtype([("int_evals", "<i4"), ("q6_evals", "<i4"), ("qbil_evals", "<i4"),
                        ("int_oob", "<i4")])

u8p = C.POINTER(C.c_uint8)
i32p = C.POINTER(C.c_int32)


class MV(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32)]


class Tiling(C.Structure):
    """Launch shape: wavefronts per call for each block size, and the tile
    workgroups (wavefronts, calls per workgroup) for the 8x8/16x16 calls
    (tile_waves 0: one workgroup per call)."""
    _fields_ = [("waves8", C.c_int32), ("waves16", C.c_int32), ("waves32", C.c_int32),
                ("waves64", C.c_int32), ("tile_waves", C.c_int32), ("tile_chunk", C.c_int32),
                ("tile_big", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TcmeError(RuntimeError):
    pass


_lib = None


def build() -> None:
    """Compile libtcme.so for gfx950 (hipcc; runs without a GPU)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "csrc")], check=True)


def _share_torch_runtime() -> None:
    """If PyTorch is importable, initialise its HIP runtime before libtcme.so is
    loaded.  torch ships its own libamdhip64 with the same soname, so loading
    torch first makes both share one runtime (device pointers and streams from
    torch are then valid for the library); loading libtcme.so first would give
    the process two HIP runtimes and torch would then find no GPU."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise TcmeError(f"{path} is not built: run __graft_entry__.build()")
    _share_torch_runtime()
    L = C.CDLL(path)
    L.tc_motion_est.restype = MV
    L.tc_motion_est.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int, C.c_int,
                                C.c_int, C.c_int, i32p]
    L.tcme_last_error.restype = C.c_char_p
    L.tcme_version.restype = C.c_char_p
    L.tcme_create.argtypes = [C.POINTER(C.c_void_p), C.c_int, C.c_int]
    L.tcme_destroy.argtypes = [C.c_void_p]
    L.tcme_destroy.restype = None
    L.tcme_pic_create.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.tcme_pic_destroy.argtypes = [C.c_void_p, C.c_int]
    L.tcme_pic_upload.argtypes = [C.c_void_p, C.c_int, u8p, C.c_int, C.c_void_p]
    L.tcme_pic_upload_device.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int, C.c_void_p]
    L.tcme_pic_luma.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_void_p), i32p]
    L.tcme_pic_finish.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
    L.tcme_search.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                              C.c_void_p]
    L.tcme_search_host.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    L.tcme_set_tiling.argtypes = [C.c_void_p, C.POINTER(Tiling)]
    L.tcme_get_tiling.argtypes = [C.c_void_p, C.POINTER(Tiling)]
    L.tcme_autotune.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int,
                                C.c_void_p, C.POINTER(Tiling), C.POINTER(C.c_float)]
    L.tcme_set_timing.argtypes = [C.c_void_p, C.c_int]
    L.tcme_pic_create_plane.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.tcme_global_mv.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
    L.tcme_global_mv_static.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p,
                                        C.c_void_p]
    L.tcme_tick_search.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64,
                                   C.c_void_p, C.c_void_p]
    L.tcme_tick_layout.restype = C.c_int64
    L.tcme_tick_layout.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p,
                                   C.c_int64]
    L.tcme_search_chain.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_int64, C.c_void_p]
    L.tcme_sad.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                           C.c_void_p]
    for f in (L.tcme_predict_luma, L.tcme_predict_chroma):
        f.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_int64, C.c_void_p,
                      C.c_void_p]
    L.tcme_last_timing.argtypes = [C.c_void_p, C.POINTER(C.c_float), C.c_int]
    L.tcme_graph_create.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.POINTER(C.c_void_p)]
    L.tcme_graph_launch.argtypes = [C.c_void_p, C.c_void_p]
    L.tcme_graph_destroy.argtypes = [C.c_void_p]
    L.tcme_graph_destroy.restype = None
    _lib = L
    return L


def _check(rc: int) -> int:
    if rc < 0:
        raise TcmeError(f"tcme error {rc}: {_lib.tcme_last_error().decode()}")
    return rc


def motion_est(ref: np.ndarray, ref_stride: int, ref_w: int, ref_h: int, orig: np.ndarray,
               orig_off: int, orig_stride: int, center_x: int, center_y: int, blk_size: int,
               search_range: int):
    """The reference call through the drop-in symbol.

    ``ref`` and ``orig`` are uint8 buffers (any shape, C-contiguous); ``orig_off``
    is the byte offset of the block's top-left sample inside ``orig``.
    Returns ``((mv_x, mv_y), sad)`` with mv in absolute quarter-pel units.
    """
    L = load()
    ref = np.ascontiguousarray(ref, np.uint8)
    orig = np.ascontiguousarray(orig, np.uint8)
    sad = C.c_int32()
    mv = L.tc_motion_est(ref.ctypes.data_as(u8p), int(ref_stride), int(ref_w), int(ref_h),
                         C.cast(orig.ctypes.data + int(orig_off), u8p), int(orig_stride),
                         int(center_x), int(center_y), int(blk_size), int(search_range),
                         C.byref(sad))
    return (mv.x, mv.y), sad.value


class Engine:
    """A tcme context: device pictures and batched searches."""

    def __init__(self, device: int = 0, max_pics: int = 1024):
        self.L = load()
        h = C.c_void_p()
        _check(self.L.tcme_create(C.byref(h), device, max_pics))
        self.h = h
        self.pics: dict[int, tuple[int, int]] = {}

    def close(self) -> None:
        if self.h:
            self.L.tcme_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def pic(self, plane: np.ndarray, stream: int = 0) -> int:
        """New device picture holding a 2-D uint8 plane (uploaded + half-sample planes)."""
        plane = np.ascontiguousarray(plane, np.uint8)
        hh, ww = plane.shape
        pid = _check(self.L.tcme_pic_create(self.h, ww, hh))
        self.pics[pid] = (ww, hh)
        self.upload(pid, plane, stream)
        return pid

    def new_pic(self, w: int, h: int) -> int:
        pid = _check(self.L.tcme_pic_create(self.h, w, h))
        self.pics[pid] = (w, h)
        return pid

    def plane(self, plane: np.ndarray, stream: int = 0) -> int:
        """New plain picture (no anchor plane) holding a 2-D uint8 plane."""
        plane = np.ascontiguousarray(plane, np.uint8)
        hh, ww = plane.shape
        pid = _check(self.L.tcme_pic_create_plane(self.h, ww, hh))
        self.pics[pid] = (ww, hh)
        self.upload(pid, plane, stream)
        return pid

    def predict_device(self, preds_ptr: int, n: int, out_ptr: int, out_bytes: int,
                       status_ptr: int | None = None, chroma: bool = False, stream: int = 0):
        """Device descriptors (PRED_DTYPE) -> predicted blocks in a device
        buffer; asynchronous on ``stream``."""
        f = self.L.tcme_predict_chroma if chroma else self.L.tcme_predict_luma
        _check(f(self.h, C.c_void_p(preds_ptr), n, C.c_void_p(out_ptr), int(out_bytes),
                 C.c_void_p(status_ptr) if status_ptr else None, C.c_void_p(stream)))

    def global_mv(self, pairs) -> np.ndarray:
        """[(cur_pic, ref_pic)] -> int32 [n, 3] = (glob_mv_x, glob_mv_y, status)."""
        import torch
        p = torch.tensor(np.asarray(pairs, np.int32).reshape(-1, 2)).cuda()
        out = torch.zeros((p.shape[0], 3), dtype=torch.int32, device="cuda")
        _check(self.L.tcme_global_mv(self.h, C.c_void_p(p.data_ptr()), p.shape[0],
                                     C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def sad(self, descs: np.ndarray, buf: np.ndarray | None = None) -> np.ndarray:
        """Host SAD descriptors (SAD_DTYPE) [+ host buffer] -> int32 SADs."""
        import torch
        descs = np.ascontiguousarray(descs, SAD_DTYPE)
        d = torch.from_numpy(descs.view(np.uint8).copy()).cuda()
        b = torch.from_numpy(np.ascontiguousarray(buf, np.uint8)).cuda() if buf is not None else None
        out = torch.zeros(len(descs), dtype=torch.int32, device="cuda")
        _check(self.L.tcme_sad(self.h, C.c_void_p(d.data_ptr()), len(descs),
                               C.c_void_p(b.data_ptr()) if b is not None else None,
                               int(b.numel()) if b is not None else 0,
                               C.c_void_p(out.data_ptr()), None))
        torch.cuda.synchronize()
        return out.cpu().numpy()

    def predict(self, preds: np.ndarray, chroma: bool = False):
        """Host descriptors -> (uint8 output buffer, status per block)."""
        import torch
        preds = np.ascontiguousarray(preds, PRED_DTYPE)
        end = preds["out_off"] + (preds["blk"] - 1).astype(np.int64) * preds["out_stride"] + \
            preds["blk"]
        nbytes = int(max(end.max(), 1))
        d_p = torch.from_numpy(preds.view(np.uint8).copy()).cuda()
        d_o = torch.zeros(nbytes, dtype=torch.uint8, device="cuda")
        d_s = torch.zeros(len(preds), dtype=torch.int32, device="cuda")
        self.predict_device(d_p.data_ptr(), len(preds), d_o.data_ptr(), nbytes, d_s.data_ptr(),
                            chroma)
        torch.cuda.synchronize()
        return d_o.cpu().numpy(), d_s.cpu().numpy()

    def upload(self, pid: int, plane: np.ndarray, stream: int = 0) -> None:
        plane = np.ascontiguousarray(plane, np.uint8)
        _check(self.L.tcme_pic_upload(self.h, pid, plane.ctypes.data_as(u8p), plane.strides[0],
                                      C.c_void_p(stream)))

    def upload_device(self, pid: int, dev_ptr: int, stride: int, stream: int = 0) -> None:
        _check(self.L.tcme_pic_upload_device(self.h, pid, C.c_void_p(dev_ptr), stride,
                                             C.c_void_p(stream)))

    def free(self, pid: int) -> None:
        _check(self.L.tcme_pic_destroy(self.h, pid))
        self.pics.pop(pid, None)

    def search(self, calls: np.ndarray, stats: bool = False, stream: int = 0):
        """Host call list (CALL_DTYPE or int32 [N,8]) -> results (and stats)."""
        calls = np.ascontiguousarray(calls)
        if calls.dtype != CALL_DTYPE:
            calls = np.ascontiguousarray(calls, np.int32).reshape(-1, 8).view(CALL_DTYPE).ravel()
        n = calls.shape[0]
        res = np.zeros(n, RESULT_DTYPE)
        st = np.zeros(n, STATS_DTYPE) if stats else None
        _check(self.L.tcme_search_host(self.h, calls.ctypes.data, n, res.ctypes.data,
                                       st.ctypes.data if stats else None, C.c_void_p(stream)))
        return (res, st) if stats else res

    def tick_search(self, recs: np.ndarray, res_ptr: int, res_cap: int, stream: int = 0):
        """One batched time step (tcme_tick_search) over host tcme_tick_pic
        records into a device result buffer; returns the offsets (n + 1)."""
        recs = np.ascontiguousarray(recs, TICK_PIC_DTYPE)
        offs = np.zeros(len(recs) + 1, np.int64)
        _check(self.L.tcme_tick_search(self.h, recs.ctypes.data, len(recs), C.c_void_p(res_ptr),
                                       int(res_cap), offs.ctypes.data, C.c_void_p(stream)))
        return offs

    def chain_ctus(self, pid: int) -> int:
        w, h = self.pics[pid]
        return ((w + 63) // 64) * ((h + 63) // 64)

    def chain_device(self, frames: np.ndarray, calls_ptr: int, counts_ptr: int, ctu_cap: int,
                     stream: int = 0) -> None:
        """The predictor chain over host frame records (CHAIN_FRAME_DTYPE) into
        device buffers; asynchronous on ``stream``."""
        frames = np.ascontiguousarray(frames, CHAIN_FRAME_DTYPE)
        _check(self.L.tcme_search_chain(self.h, frames.ctypes.data, len(frames),
                                        C.c_void_p(calls_ptr), C.c_void_p(counts_ptr),
                                        int(ctu_cap), C.c_void_p(stream)))

    def chain(self, frames):
        """frames: [(cur_pic, [ref_pic, ...], sr, lambda)] -> per frame an int32
        [N, 10] array of its searches (ctu, ref, px, py, cx, cy, n, mv_x, mv_y,
        sad), CTUs in raster order, each CTU's in the order they were made."""
        import torch
        rec = np.zeros(len(frames), CHAIN_FRAME_DTYPE)
        for i, (cur, refs, sr, lam) in enumerate(frames):
            rec[i]["cur_pic"], rec[i]["nrefs"] = cur, len(refs)
            rec[i]["ref_pic"][:len(refs)] = refs
            rec[i]["sr"], rec[i]["lambda"] = sr, lam
        nct = [self.chain_ctus(int(f[0])) for f in frames]
        total = sum(nct)
        d_calls = torch.empty(total * CHAIN_MAX_CALLS * CHAIN_CALL_DTYPE.itemsize,
                              dtype=torch.uint8, device="cuda")
        d_counts = torch.full((total,), -1, dtype=torch.int32, device="cuda")
        self.chain_device(rec, d_calls.data_ptr(), d_counts.data_ptr(), total)
        torch.cuda.synchronize()
        counts = d_counts.cpu().numpy()
        allc = d_calls.cpu().numpy().view(np.int32).reshape(total, CHAIN_MAX_CALLS, 10)
        out, k0 = [], 0
        for n in nct:
            rows = []
            for k in range(n):
                c = int(counts[k0 + k])
                assert 0 <= c <= CHAIN_MAX_CALLS, c
                r = allc[k0 + k, :c, :9]
                rows.append(np.concatenate([np.full((c, 1), k, np.int32), r], 1))
            out.append(np.concatenate(rows) if rows else np.zeros((0, 10), np.int32))
            k0 += n
        return out

    def tiling(self) -> Tiling:
        t = Tiling()
        _check(self.L.tcme_get_tiling(self.h, C.byref(t)))
        return t

    def set_tiling(self, waves8: int, waves16: int, waves32: int, waves64: int,
                   tile_waves: int = 8, tile_chunk: int = 85, tile_big: int = 1) -> None:
        t = Tiling(waves8, waves16, waves32, waves64, tile_waves, tile_chunk, tile_big)
        _check(self.L.tcme_set_tiling(self.h, C.byref(t)))

    def autotune(self, calls_ptr: int, n: int, res_ptr: int, reps: int = 3, stream: int = 0):
        """Picks the fastest launch shape for this device call list; returns
        (Tiling, ms per launch)."""
        t = Tiling()
        ms = C.c_float()
        _check(self.L.tcme_autotune(self.h, C.c_void_p(calls_ptr), n, C.c_void_p(res_ptr), reps,
                                    C.c_void_p(stream), C.byref(t), C.byref(ms)))
        return t, ms.value

    def set_timing(self, on: bool) -> None:
        _check(self.L.tcme_set_timing(self.h, int(on)))

    def last_timing(self) -> list[float]:
        """Per-kernel ms of the last dispatch (waits for it), in launch order."""
        ms = (C.c_float * 4)()
        n = _check(self.L.tcme_last_timing(self.h, ms, 4))
        return [ms[k] for k in range(n)]

    def search_device(self, calls_ptr: int, n: int, res_ptr: int, stats_ptr: int | None = None,
                      stream: int = 0) -> None:
        """Device call list -> device results; asynchronous on ``stream``."""
        _check(self.L.tcme_search(self.h, C.c_void_p(calls_ptr), n, C.c_void_p(res_ptr),
                                  C.c_void_p(stats_ptr) if stats_ptr else None,
                                  C.c_void_p(stream)))

    def graph(self, calls_ptr: int, n: int, res_ptr: int,
              stats_ptr: int | None = None) -> "Graph":
        """The search of this device call list captured as a HIP graph."""
        g = C.c_void_p()
        _check(self.L.tcme_graph_create(self.h, C.c_void_p(calls_ptr), n, C.c_void_p(res_ptr),
                                        C.c_void_p(stats_ptr) if stats_ptr else None,
                                        C.byref(g)))
        return Graph(self, g)


class Graph:
    """A captured search (tcme_graph_*): replay with ``launch(stream)``."""

    def __init__(self, eng: Engine, handle: C.c_void_p):
        self.eng, self.h = eng, handle

    def launch(self, stream: int = 0) -> None:
        _check(self.eng.L.tcme_graph_launch(self.h, C.c_void_p(stream)))

    def close(self) -> None:
        if self.h:
            self.eng.L.tcme_graph_destroy(self.h)
            self.h = None


def tick_layout(w: int, h: int, nrefs: int, sr: int) -> np.ndarray:
    """A picture shape's call list as tcme_tick_search lays it out (CALL_DTYPE,
    ref_pic = reference index, cur_pic = -1); host only, no GPU needed."""
    L = load()
    n = _check(L.tcme_tick_layout(None, w, h, nrefs, sr, None, 0))
    out = np.zeros(n, CALL_DTYPE)
    _check(L.tcme_tick_layout(None, w, h, nrefs, sr, out.ctypes.data, n))
    return out


def make_calls(ref_pic, cur_pic, px, py, cx, cy, blk, sr) -> np.ndarray:
    """Vectorised CALL_DTYPE constructor (scalars broadcast)."""
    cols = np.broadcast_arrays(*[np.asarray(v, np.int32) for v in
                                 (ref_pic, cur_pic, px, py, cx, cy, blk, sr)])
    out = np.zeros(cols[0].size, CALL_DTYPE)
    for name, col in zip(CALL_DTYPE.names, cols):
        out[name] = col.ravel()
    return out
