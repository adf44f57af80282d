"""ctypes access to the checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py``.  The product never imports this package.

* ``Oracle``    - oracle/_build/libtcme_oracle.so, the C restatement
* ``Reference`` - oracle/_ref/libtcref.so, the reference codec compiled from
  its own sources (built only where the reference tree is present)
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "_build", "libtcme_oracle.so")
REF_DIR = os.path.join(HERE, "_ref")
REF_SO = os.path.join(REF_DIR, "libtcref.so")
REFERENCE_TREE = os.environ.get("TCME_REFERENCE", "/root/reference")

u8p = C.POINTER(C.c_uint8)
i16p = C.POINTER(C.c_int16)
i32p = C.POINTER(C.c_int32)


class MV(C.Structure):
    _fields_ = [("x", C.c_int32), ("y", C.c_int32)]


class Stats(C.Structure):
    _fields_ = [("int_evals", C.c_int32), ("int_oob", C.c_int32),
                ("q6_evals", C.c_int32), ("qbil_evals", C.c_int32),
                ("q_skip", C.c_int32)]


def ptr(a: np.ndarray, typ=u8p):
    """Data pointer of `a`.  The wrappers pass ``a.strides[0]`` as the row
    pitch, so a 2-D plane must be row-major with contiguous rows (a
    column-major array, e.g. what ``np.roll`` over both axes can return,
    would be read transposed)."""
    if a.ndim >= 2 and (a.strides[-1] != a.itemsize or a.strides[0] < a.shape[-1] * a.itemsize):
        raise ValueError(f"plane must be row-major with contiguous rows, strides {a.strides}")
    return a.ctypes.data_as(typ)


def build(with_reference: bool = True) -> None:
    """Compile the restatement and, when the reference tree exists, oracle/_ref."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if with_reference and os.path.isdir(os.path.join(REFERENCE_TREE, "src")):
        subprocess.run(["make", "-s", "-j8", "-C", HERE, "ref", f"REF={REFERENCE_TREE}"],
                       check=True)


def reference_available() -> bool:
    return os.path.exists(REF_SO)


class Oracle:
    def __init__(self, path: str = ORACLE_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (run oracle.build())")
        L = self.lib = C.CDLL(path)
        L.orc_sad.restype = C.c_int32
        L.orc_sad.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int]
        for f in ("orc_half_h", "orc_half_v", "orc_half_hv"):
            getattr(L, f).restype = C.c_int
            getattr(L, f).argtypes = [u8p, C.c_int, C.c_int, C.c_int]
        L.orc_qpel6.restype = C.c_int
        L.orc_qpel6.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.orc_bilin.restype = C.c_int
        L.orc_bilin.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        for f in ("orc_sad_subpel", "orc_sad_qpel6"):
            getattr(L, f).restype = C.c_int32
            getattr(L, f).argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.orc_motion_est.restype = MV
        L.orc_motion_est.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int,
                                     C.c_int, C.c_int, C.c_int, i32p, C.POINTER(Stats)]
        L.orc_motion_est_batch.restype = None
        L.orc_motion_est_batch.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int,
                                           i32p, C.c_int, i32p, C.POINTER(Stats)]
        L.orc_algo_bytes.restype = C.c_uint64
        L.orc_algo_bytes.argtypes = [C.POINTER(Stats), C.c_int]
        L.orc_luma_px.restype = C.c_int
        L.orc_luma_px.argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_int]
        for f in ("orc_inter_predict", "orc_inter_predict_chroma"):
            getattr(L, f).restype = None
            getattr(L, f).argtypes = [u8p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, u8p,
                                      C.c_int, C.c_int]
        L.orc_mvp.restype = None
        L.orc_mvp.argtypes = [C.c_int, C.c_int, i16p, i16p, u8p, i32p, i32p]
        L.orc_global_mv.restype = None
        L.orc_global_mv.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.orc_se_bits.restype = C.c_int
        L.orc_se_bits.argtypes = [C.c_int32]
        L.orc_chain_frame.restype = C.c_int
        L.orc_chain_frame.argtypes = [C.POINTER(u8p), C.c_int, C.c_int, u8p, C.c_int, C.c_int,
                                      C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p, i32p,
                                      C.c_int]

    # -- convenience wrappers over numpy planes (C-contiguous uint8 2-D) --
    def motion_est(self, ref, cur, px, py, cx, cy, n, sr, stats=False):
        h, w = ref.shape
        sad = C.c_int32()
        st = Stats()
        o = cur[py:, px:]
        mv = self.lib.orc_motion_est(ptr(ref), ref.strides[0], w, h,
                                     C.cast(cur.ctypes.data + int(py) * cur.strides[0] + int(px), u8p),
                                     cur.strides[0], cx, cy, n, sr, C.byref(sad), C.byref(st))
        del o
        if stats:
            return (mv.x, mv.y, sad.value), st
        return mv.x, mv.y, sad.value

    def motion_est_batch(self, ref, cur, calls, want_stats=False):
        """calls: int32 [N,6] = px, py, cx, cy, n, sr -> int32 [N,3] (mv_x, mv_y, sad)."""
        calls = np.ascontiguousarray(calls, dtype=np.int32)
        n = calls.shape[0]
        out = np.zeros((n, 3), np.int32)
        st = (Stats * n)() if want_stats else None
        h, w = ref.shape
        self.lib.orc_motion_est_batch(ptr(ref), ref.strides[0], w, h, ptr(cur), cur.strides[0],
                                      ptr(calls, i32p), n, ptr(out, i32p), st)
        if want_stats:
            sarr = np.frombuffer(st, dtype=np.int32).reshape(n, 5).copy()
            return out, sarr
        return out

    def algo_bytes(self, stats_row, n):
        st = Stats(*[int(v) for v in stats_row])
        return int(self.lib.orc_algo_bytes(C.byref(st), n))

    def inter_predict(self, ref, mvx, mvy, n):
        h, w = ref.shape
        dst = np.zeros((n, n), np.uint8)
        self.lib.orc_inter_predict(ptr(ref), ref.strides[0], w, h, mvx, mvy, ptr(dst), n, n)
        return dst

    def inter_predict_chroma(self, ref, mvx, mvy, n):
        h, w = ref.shape
        dst = np.zeros((n, n), np.uint8)
        self.lib.orc_inter_predict_chroma(ptr(ref), ref.strides[0], w, h, mvx, mvy, ptr(dst),
                                          n, n)
        return dst

    def inter_predict_list(self, ref, mvs, sizes, offs, out, chroma=False):
        """Every block k (MV mvs[k], size sizes[k], packed at out[offs[k]:]
        with row stride sizes[k]) into the 1-D uint8 array `out`."""
        f = self.lib.orc_inter_predict_chroma if chroma else self.lib.orc_inter_predict
        h, w = ref.shape
        rp, st, base = ptr(ref), ref.strides[0], out.ctypes.data
        assert out.dtype == np.uint8 and out.ndim == 1 and out.flags.c_contiguous
        for k in range(len(sizes)):
            n, o = int(sizes[k]), int(offs[k])
            assert 0 <= o and o + n * n <= out.size
            f(rp, st, w, h, int(mvs[k][0]), int(mvs[k][1]), C.cast(base + o, u8p), n, n)
        return out

    def mvp(self, cx, cy, dx, dy, intra):
        ox, oy = C.c_int32(), C.c_int32()
        dx = np.ascontiguousarray(dx, np.int16)
        dy = np.ascontiguousarray(dy, np.int16)
        intra = np.ascontiguousarray(intra, np.uint8)
        self.lib.orc_mvp(cx, cy, ptr(dx, i16p), ptr(dy, i16p), ptr(intra), C.byref(ox),
                         C.byref(oy))
        return ox.value, oy.value

    CHAIN_ROW = 10

    def chain_frame(self, refs, cur, sr, lam, mvp=None, me=None):
        """The ME-only predictor chain over one frame: int32 [N, 10] rows
        (ctu, ref, px, py, cx, cy, n, mv_x, mv_y, sad) in call order.  mvp /
        me: None (the restatement) or C function pointers (e.g. the
        reference's own, from Reference.fn_pointers())."""
        h, w = cur.shape
        refs = [np.ascontiguousarray(r, np.uint8) for r in refs]
        assert all(r.shape == cur.shape and r.strides == refs[0].strides for r in refs)
        arr = (u8p * len(refs))(*[ptr(r) for r in refs])
        cols, rows = (w + 63) // 64, (h + 63) // 64
        cap = cols * rows * 169 * len(refs)
        out = np.zeros((cap, self.CHAIN_ROW), np.int32)
        n = self.lib.orc_chain_frame(arr, len(refs), refs[0].strides[0], ptr(cur), cur.strides[0],
                                     w, h, sr, lam, mvp, me, ptr(out, i32p), cap)
        assert n <= cap
        return out[:n].copy()

    def global_mv(self, cur, ref):
        h, w = ref.shape
        gx, gy = C.c_int32(), C.c_int32()
        self.lib.orc_global_mv(ptr(cur), cur.strides[0], ptr(ref), ref.strides[0], w, h,
                               C.byref(gx), C.byref(gy))
        return gx.value, gy.value


class Reference:
    """The reference's own functions, compiled from its sources."""

    def __init__(self, path: str = REF_SO):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} not built (reference tree absent?)")
        L = self.lib = C.CDLL(path)
        L.tc_sad.restype = C.c_int32
        L.tc_sad.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int]
        L.tc_sad_subpel.restype = C.c_int32
        L.tc_sad_subpel.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int]
        L.tc_motion_est.restype = MV
        L.tc_motion_est.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, C.c_int,
                                    C.c_int, C.c_int, C.c_int, i32p]
        L.tc_inter_predict.restype = None
        L.tc_inter_predict.argtypes = [u8p, C.c_int, C.c_int, C.c_int, MV, u8p, C.c_int,
                                       C.c_int]
        L.tc_inter_predict_chroma.restype = None
        L.tc_inter_predict_chroma.argtypes = [u8p, C.c_int, C.c_int, C.c_int, MV, u8p,
                                              C.c_int, C.c_int]
        L.refx_qt_mvp.restype = C.c_int
        L.refx_qt_mvp.argtypes = [C.c_int, C.c_int, i16p, i16p, u8p, i32p, i32p]
        L.refx_global_mv.restype = C.c_int
        L.refx_global_mv.argtypes = [u8p, C.c_int, u8p, C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.refx_qt_mvp_v.restype = None
        L.refx_motion_est_batch.restype = C.c_int
        L.refx_motion_est_batch.argtypes = [u8p, C.c_int, C.c_int, C.c_int, u8p, C.c_int, i32p,
                                            C.c_int, i32p]

    def fn_pointers(self):
        """(qt_mvp, tc_motion_est) of the reference as C function pointers, for
        Oracle.chain_frame."""
        return (C.cast(self.lib.refx_qt_mvp_v, C.c_void_p).value,
                C.cast(self.lib.tc_motion_est, C.c_void_p).value)

    def motion_est(self, ref, cur, px, py, cx, cy, n, sr, w=None, h=None):
        hh, ww = ref.shape
        w = ww if w is None else w
        h = hh if h is None else h
        sad = C.c_int32()
        mv = self.lib.tc_motion_est(ptr(ref), ref.strides[0], w, h,
                                    C.cast(cur.ctypes.data + int(py) * cur.strides[0] + int(px), u8p),
                                    cur.strides[0], cx, cy, n, sr, C.byref(sad))
        return mv.x, mv.y, sad.value

    def motion_est_batch(self, ref, cur, calls):
        """calls: int32 [N,6] = px, py, cx, cy, n, sr -> int32 [N,3], the
        reference's tc_motion_est for each (one foreign call: the GIL is
        released for the whole batch)."""
        calls = np.ascontiguousarray(calls, dtype=np.int32)
        out = np.zeros((calls.shape[0], 3), np.int32)
        h, w = ref.shape
        self.lib.refx_motion_est_batch(ptr(ref), ref.strides[0], w, h, ptr(cur), cur.strides[0],
                                       ptr(calls, i32p), calls.shape[0], ptr(out, i32p))
        return out

    def inter_predict(self, ref, mvx, mvy, n, w=None, h=None):
        hh, ww = ref.shape
        dst = np.zeros((n, n), np.uint8)
        self.lib.tc_inter_predict(ptr(ref), ref.strides[0], ww if w is None else w,
                                  hh if h is None else h, MV(mvx, mvy), ptr(dst), n, n)
        return dst

    def inter_predict_chroma(self, ref, mvx, mvy, n):
        hh, ww = ref.shape
        dst = np.zeros((n, n), np.uint8)
        self.lib.tc_inter_predict_chroma(ptr(ref), ref.strides[0], ww, hh, MV(mvx, mvy),
                                         ptr(dst), n, n)
        return dst

    def mvp(self, cx, cy, dx, dy, intra):
        ox, oy = C.c_int32(), C.c_int32()
        dx = np.ascontiguousarray(dx, np.int16)
        dy = np.ascontiguousarray(dy, np.int16)
        intra = np.ascontiguousarray(intra, np.uint8)
        self.lib.refx_qt_mvp(cx, cy, ptr(dx, i16p), ptr(dy, i16p), ptr(intra), C.byref(ox),
                             C.byref(oy))
        return ox.value, oy.value

    def global_mv(self, cur, ref):
        h, w = ref.shape
        gx, gy = C.c_int32(), C.c_int32()
        self.lib.refx_global_mv(ptr(cur), cur.strides[0], ptr(ref), ref.strides[0], w, h,
                                C.byref(gx), C.byref(gy))
        return gx.value, gy.value
