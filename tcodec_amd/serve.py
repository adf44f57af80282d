"""Batched and served forms of the search: many encode streams, one launch per tick.

A tick takes the next coded picture of every active stream, and searches all
their calls (every in-frame quadtree node against the picture's references, as
in ``workload.frame_calls_by_ctu``) in ONE ``tcme_search`` launch.  Streams
keep their reference pictures resident on the device: each stream owns a ring
of 5 device pictures (4 references + the current one, the reference encoder's
4-slot picture buffer, src/encoder.c:3085-3097) allocated when it is admitted
and freed when it finishes; a new picture is copied into the ring slot of the
one leaving the buffer (device to device, anchors rebuilt).

``Server`` is the serving loop: requests arrive over time (Poisson), wait in a
queue, and are admitted while their pictures fit an HBM budget and the tick's
call count stays under a work bound; admitted streams join the very next tick
(continuous batching, no draining).  A frame is ready when its stream's
previous frame is done (the encoder's dependency on reconstructed references);
its latency runs from ready to the end of the tick that searched it.

Source pictures come from device-resident clips per resolution (``Clip``), so
the timed loop does no host-side generation.  Results are pure functions of
(pictures, calls), so a stream's results are the same alone and in any mix:
``run_alone`` and tests/test_gpu_serve.py check exactly that.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from . import me, synth, workload

RES = {"240p": (320, 240), "720p": (1280, 720), "1080p": (1920, 1080), "2160p": (3840, 2160)}


def pic_bytes(w: int, h: int) -> int:
    """Device bytes of one reference picture: luma + anchors (8 B per sample)
    + tail slack, rows padded to 64."""
    pitch = (w + 63) // 64 * 64
    return pitch * h * 9 + 512


def stream_bytes(w: int, h: int) -> int:
    return 5 * pic_bytes(w, h)


class Clip:
    """F source luma planes of one resolution, resident on the device."""

    def __init__(self, w: int, h: int, frames: int, seed: int, kind: str = "scene"):
        self.w, self.h, self.frames = w, h, frames
        if kind == "scene":
            sc = synth.Scene(w, h, seed=seed, rects=48 if w >= 1920 else 16, pan=(0.75, 0.25))
            planes = [sc.frame(t)[0] for t in range(frames)]
        else:
            planes = [f[0] for f in synth.gradient_video(w, h, frames, seed=seed)]
        self.host = np.stack(planes)
        self.dev = torch.from_numpy(self.host).cuda()

    def plane_ptr(self, t: int) -> int:
        return self.dev[t % self.frames].data_ptr()


class DeviceGradientClip(Clip):
    """The diagonal-gradient source (translated (+3,-2) px per frame, uniform
    integer noise in [-4,4]) rendered on the device with the device's own
    counter RNG seeded per stream: many streams of large pictures without
    host-side generation.  Same structure as synth.gradient_video; the noise
    values differ from the host generator's."""

    def __init__(self, w: int, h: int, frames: int, seed: int):
        self.w, self.h, self.frames = w, h, frames
        dev = torch.device("cuda")
        yy = torch.arange(h, device=dev, dtype=torch.int64).view(-1, 1)
        xx = torch.arange(w, device=dev, dtype=torch.int64).view(1, -1)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        out = torch.empty((frames, h, w), dtype=torch.uint8, device=dev)
        for t in range(frames):
            g = ((xx - 3 * t) + (yy + 2 * t)) * 255 // (w + h - 2)
            nz = torch.randint(-4, 5, (h, w), generator=gen, device=dev)
            out[t] = (g + nz).clamp_(0, 255).to(torch.uint8)
        self.dev = out
        self.host = None


class DeviceSceneClip(Clip):
    """The moving-rectangle scene of ``synth.Scene`` (smooth value-noise
    background scaled to 16..235, K textured rectangles 32..256 px per side
    moving at quarter-pel velocities in [-16, 16] px/frame and wrapping at the
    frame edge, a global pan, Gaussian noise sigma 2) rendered on the device
    in float32 with the device's own RNG, so that long clips of large pictures
    (every picture distinct) cost no host time.  Same structure as the host
    generator; the values differ from it."""

    PERIODS, AMPS = (128, 48, 16), (1.0, 0.5, 0.25)

    def __init__(self, w: int, h: int, frames: int, seed: int, rects: int | None = None,
                 pan: tuple[float, float] = (0.75, 0.25), sigma: float = 2.0,
                 device: str = "cuda"):
        self.w, self.h, self.frames = w, h, frames
        self.host = None
        dev = torch.device(device)
        rects = (48 if w >= 1920 else 16) if rects is None else rects
        rng = np.random.default_rng(seed)
        gen = torch.Generator(device=dev)
        gen.manual_seed(seed)
        rw = rng.integers(32, 257, rects)
        rh = rng.integers(32, 257, rects)
        x0 = rng.random(rects) * (w + rw) - rw
        y0 = rng.random(rects) * (h + rh) - rh
        vx = rng.integers(-64, 65, rects) / 4.0
        vy = rng.integers(-64, 65, rects) / 4.0
        base = rng.integers(40, 216, rects).astype(np.float64)
        tex = [(torch.rand((int(rh[k]) // 8 + 2, int(rw[k]) // 8 + 2), generator=gen,
                           device=dev) * 2 - 1) * 40.0 for k in range(rects)]
        # one lattice per octave covering every frame's panned window
        ox = [-pan[0] * t for t in (0, frames - 1)]
        oy = [-pan[1] * t for t in (0, frames - 1)]
        lats = []
        for per in self.PERIODS:
            gx0 = int(np.floor(min(ox) / per)) - 1
            gy0 = int(np.floor(min(oy) / per)) - 1
            gx1 = int(np.floor((max(ox) + w) / per)) + 2
            gy1 = int(np.floor((max(oy) + h) / per)) + 2
            lats.append((gx0, gy0, torch.rand((gy1 - gy0 + 1, gx1 - gx0 + 1), generator=gen,
                                              device=dev)))
        xs = torch.arange(w, device=dev, dtype=torch.float32)
        ys = torch.arange(h, device=dev, dtype=torch.float32)
        out = torch.empty((frames, h, w), dtype=torch.uint8, device=dev)
        for t in range(frames):
            acc = torch.zeros((h, w), device=dev)
            for per, amp, (gx0, gy0, lat) in zip(self.PERIODS, self.AMPS, lats):
                u = (xs - pan[0] * t) / per
                v = (ys - pan[1] * t) / per
                iu, iv = torch.floor(u), torch.floor(v)
                fu, fv = u - iu, v - iv
                fu = fu * fu * (3 - 2 * fu)
                fv = fv * fv * (3 - 2 * fv)
                cu = (iu.long() - gx0)
                cv = (iv.long() - gy0)
                l0, l1 = lat[cv], lat[cv + 1]
                top = l0[:, cu] + (l0[:, cu + 1] - l0[:, cu]) * fu[None, :]
                bot = l1[:, cu] + (l1[:, cu + 1] - l1[:, cu]) * fu[None, :]
                acc += amp * (top + (bot - top) * fv[:, None])
            img = 16.0 + acc / sum(self.AMPS) * (235.0 - 16.0)
            for k in range(rects):
                kw, kh = int(rw[k]), int(rh[k])
                px = (x0[k] + vx[k] * t + kw) % (w + kw) - kw
                py = (y0[k] + vy[k] * t + kh) % (h + kh) - kh
                xa, xb = max(0, int(np.ceil(px))), min(w, int(np.ceil(px + kw)))
                ya, yb = max(0, int(np.ceil(py))), min(h, int(np.ceil(py + kh)))
                if xa >= xb or ya >= yb:
                    continue
                T = tex[k]
                u = (xs[xa:xb] - float(px)) / 8.0
                v = (ys[ya:yb] - float(py)) / 8.0
                iu = torch.clamp(torch.floor(u).long(), max=T.shape[1] - 2)
                iv = torch.clamp(torch.floor(v).long(), max=T.shape[0] - 2)
                fu = (u - iu)[None, :]
                fv = (v - iv)[:, None]
                t0, t1 = T[iv], T[iv + 1]
                top = t0[:, iu] + (t0[:, iu + 1] - t0[:, iu]) * fu
                bot = t1[:, iu] + (t1[:, iu + 1] - t1[:, iu]) * fu
                img[ya:yb, xa:xb] = float(base[k]) + top + (bot - top) * fv
            nz = torch.randn((h, w), generator=gen, device=dev) * sigma
            out[t] = torch.round(img + nz).clamp_(0, 255).to(torch.uint8)
        self.dev = out


def frame_template(w: int, h: int, refs: int, sr: int) -> np.ndarray:
    """A picture's call list with picture ids 0..refs-1 for the references
    and `refs` for the current picture: every quadtree node against every
    reference, in 85-call groups, reference-major."""
    return workload.frame_calls_by_ctu(w, h, list(range(refs)), refs, sr, ref_major=True)


class Templates:
    """Device call-list templates per (w, h, refs, sr) (``frame_template``)."""

    def __init__(self):
        self.cache: dict[tuple, torch.Tensor] = {}

    def get(self, w, h, refs, sr) -> torch.Tensor:
        key = (w, h, refs, sr)
        t = self.cache.get(key)
        if t is None:
            calls = frame_template(w, h, refs, sr)
            t = torch.from_numpy(calls.view(np.int32).reshape(-1, 8).copy()).cuda()
            self.cache[key] = t
        return t


@dataclass
class Stream:
    sid: int
    w: int
    h: int
    nframes: int
    gop: str                      # "P" or "B" (GOP4 ladder)
    clip: Clip
    offset: int = 0
    arrival: float = 0.0
    order: list = field(default_factory=list)
    pos: int = 0                  # next index in coding order
    ring: list = field(default_factory=list)   # device picture ids
    slot_of: dict = field(default_factory=dict)  # poc -> ring index
    ready_at: float = 0.0
    busy: bool = False            # a picture of this stream is in flight
    latencies: list = field(default_factory=list)
    results: list = field(default_factory=list)   # keep_results=True: every picture's results
    digests: list = field(default_factory=list)   # keep_results="digest": one per picture
    keep_at: set = field(default_factory=set)     # ... and these coded positions in full
    kept: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.gop == "B":
            n = 4 * max(1, (self.nframes - 1 + 3) // 4) + 1
            self.order = workload.gop4_coding_order(n)
        else:
            self.order = workload.p_coding_order(self.nframes)

    @property
    def done(self) -> bool:
        return self.pos >= len(self.order)

    def activate(self, eng: me.Engine) -> None:
        self.ring = [eng.new_pic(self.w, self.h) for _ in range(5)]
        self.slot_of = {}

    def release(self, eng: me.Engine) -> None:
        for p in self.ring:
            eng.free(p)
        self.ring = []

    def prepare(self, eng: me.Engine, stream_ptr: int, trace=None):
        """Loads the next coded picture into the ring (a device copy, anchors
        rebuilt, on stream_ptr) and returns its search as (template key,
        picture ids of the references then of the current picture), or None
        for an intra picture.  Advances pos.  With a trace.TraceWriter, the
        picture is recorded as a plane (its device picture id is the plane
        id)."""
        poc, kind, dpb = self.order[self.pos]
        self.pos += 1
        src = (self.offset + poc) % self.clip.frames
        assert all((self.offset + p) % self.clip.frames != src for p in dpb), \
            "a reference would be the same source picture as the current one"
        keep = set(dpb)
        slot = next(i for i in range(5) if i not in
                    {s for p, s in self.slot_of.items() if p in keep})
        self.slot_of = {p: s for p, s in self.slot_of.items() if p in keep}
        self.slot_of[poc] = slot
        eng.upload_device(self.ring[slot], self.clip.plane_ptr(self.offset + poc), self.w,
                          stream_ptr)
        if trace is not None:
            host = self.clip.host[src] if self.clip.host is not None else \
                self.clip.dev[src].cpu().numpy()
            trace.plane(self.ring[slot], host)
        if kind == "I" or not dpb:
            return None
        sr = workload.SR_B if kind == "B" else workload.SR_P
        return (self.w, self.h, len(dpb), sr), [self.ring[self.slot_of[p]] for p in dpb] + \
            [self.ring[slot]]


def expand_calls(tpl: Templates, items) -> tuple[torch.Tensor, list[int]]:
    """The call lists of a tick's pictures, concatenated in item order, built
    on the device with one gather per template: items are (template key,
    picture-id table) as Stream.prepare returns them.  Returns the int32
    [N, 8] list and each picture's call count."""
    groups: dict = {}
    for i, (key, lut) in enumerate(items):
        groups.setdefault(key, []).append(i)
    parts: list = [None] * len(items)
    for key, idx in groups.items():
        t = tpl.get(*key)
        luts = torch.tensor([items[i][1] for i in idx], dtype=torch.int32).cuda(non_blocking=True)
        k, n = len(idx), t.shape[0]
        calls = t.unsqueeze(0).expand(k, n, 8).clone()
        calls[:, :, 0] = torch.gather(luts, 1, t[:, 0].long().unsqueeze(0).expand(k, n))
        calls[:, :, 1] = luts[:, key[2]].unsqueeze(1)
        for j, i in enumerate(idx):
            parts[i] = calls[j]
    sizes = [p.shape[0] for p in parts]
    return (torch.cat(parts) if len(parts) > 1 else parts[0]), sizes


def tick_records(items) -> np.ndarray:
    """tcme_tick_pic records of a tick's pictures (items as Stream.prepare
    returns them)."""
    rec = np.zeros(len(items), me.TICK_PIC_DTYPE)
    for i, ((w, h, nrefs, sr), lut) in enumerate(items):
        rec[i]["cur_pic"], rec[i]["nrefs"], rec[i]["sr"] = lut[-1], nrefs, sr
        rec[i]["ref_pic"][:nrefs] = lut[:-1]
    return rec


def digests(res: torch.Tensor, sizes: list[int]) -> torch.Tensor:
    """One int64 digest per picture of a tick's results (int32 [N, 4], the
    pictures' calls concatenated): a sum of every field weighted by the
    call's position inside its picture, on the device (wrapping int64).
    Equal digests of the same picture served two ways are what the
    alone-versus-mix checks compare when keeping every result would not fit."""
    dev = res.device
    r = res.long()
    v = r[:, 0] * 1000003 + r[:, 1] * 999983 + r[:, 2] * 97 + r[:, 3]
    sz = torch.tensor(sizes, dtype=torch.int64, device=dev)
    starts = torch.cumsum(sz, 0) - sz
    local = torch.arange(r.shape[0], device=dev) - torch.repeat_interleave(starts, sz)
    cs = torch.cumsum(v * ((local * 2654435761) % 2147483647 + 1), 0)
    ends = cs[torch.cumsum(sz, 0) - 1]
    return ends - torch.cat([ends.new_zeros(1), ends[:-1]])


class Server:
    """Continuous-batching request loop over one tcme context."""

    def __init__(self, eng: me.Engine, hbm_budget: int, max_tick_calls: int = 6_000_000,
                 keep_results: bool = False, per_stream_launch: bool = False, trace=None,
                 pipeline: bool = False, gpu_timing: bool = False):
        self.eng = eng
        self.trace = trace        # trace.TraceWriter: record planes and searches
        self.per_stream = per_stream_launch
        self.budget = hbm_budget
        self.max_tick_calls = max_tick_calls
        self.keep = keep_results  # False, True (every result to the host) or "digest"
        self.tpl = Templates()
        self._stream = None
        self.queue: list[Stream] = []
        self.active: list[Stream] = []
        self.finished: list[Stream] = []
        self.dropped: list[Stream] = []   # requests larger than the whole budget
        self.used = 0
        self.frames = 0           # searched (non-intra) pictures
        self.calls = 0
        self.ticks = 0
        self.tick_ms: list[float] = []
        self.pipeline = pipeline
        self.gpu_timing = gpu_timing  # events around each launch: gpu_ms
        self.gpu_ms = 0.0             # device time of the launches' work
        self.inflight: list = []
        self.max_active = 0           # most streams searched by one launch
        self.admitted_while_busy = 0  # admissions joining streams already running

    @property
    def stream(self):
        if self._stream is None:
            self._stream = torch.cuda.current_stream()
        return self._stream

    @property
    def sptr(self) -> int:
        return self.stream.cuda_stream

    def submit(self, s: Stream) -> None:
        self.queue.append(s)

    _layout_n: dict = {}

    @classmethod
    def layout_size(cls, key) -> int:
        n = cls._layout_n.get(key)
        if n is None:
            n = cls._layout_n[key] = len(frame_template(key[0], key[1], key[2], key[3]))
        return n

    @staticmethod
    def host_calls(key, lut) -> np.ndarray:
        """A picture's call list on the host (as the library lays it out)."""
        calls = frame_template(*key).copy()
        calls["ref_pic"] = np.asarray(lut[:-1], np.int32)[calls["ref_pic"]]
        calls["cur_pic"] = lut[-1]
        return calls

    _nodes: dict = {}

    @classmethod
    def calls_per_frame(cls, s: Stream) -> int:
        """Upper bound (4 references) of a picture's calls."""
        n = cls._nodes.get((s.w, s.h))
        if n is None:
            n = cls._nodes[(s.w, s.h)] = len(workload.qt_nodes(s.w, s.h))
        return 4 * n

    def _admit(self, now: float) -> None:
        work = sum(self.calls_per_frame(s) for s in self.active)
        while self.queue and self.queue[0].arrival <= now:
            s = self.queue[0]
            need = stream_bytes(s.w, s.h)
            cost = self.calls_per_frame(s)
            if need > self.budget:
                # could never be admitted, not even by an idle server: refuse
                # it rather than let it block every request behind it
                self.queue.pop(0)
                self.dropped.append(s)
                continue
            if self.used + need > self.budget or (self.active and work + cost > self.max_tick_calls):
                break
            self.queue.pop(0)
            if self.active:
                self.admitted_while_busy += 1
            s.activate(self.eng)
            s.ready_at = now
            self.used += need
            work += cost
            self.active.append(s)
            self.max_active = max(self.max_active, len(self.active))

    def _launch(self, now: float, streams: list) -> dict:
        """Enqueues one launch over the next picture of each of `streams`
        (uploads, anchors and the search, on the server's stream) and returns
        what completing it needs."""
        items, owners = [], []
        for s in streams:
            pos = s.pos
            c = s.prepare(self.eng, self.sptr, self.trace)
            s.busy = True
            if c is not None:
                items.append(c)
                owners.append((s, pos))
        t = {"now": now, "streams": streams, "items": items, "owners": owners, "res": None,
             "sizes": [], "dig": None, "ev": None}
        if self.gpu_timing:
            t["ev"] = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
            t["ev"][0].record(self.stream)
        if items:
            # the step's call list is built on the device by the library
            # (tcme_tick_search) from per-shape layouts
            recs = tick_records(items)
            sizes = [self.layout_size(key) for key, _ in items]
            res = torch.empty((sum(sizes), 4), dtype=torch.int32, device="cuda")
            if self.per_stream:
                # the plain form: one launch per picture, back to back
                o = 0
                for k, n in enumerate(sizes):
                    self.eng.tick_search(recs[k:k + 1], res[o].data_ptr(), n, self.sptr)
                    o += n
            else:
                self.eng.tick_search(recs, res.data_ptr(), res.shape[0], self.sptr)
            t["res"], t["sizes"] = res, sizes
            t["dig"] = digests(res, sizes) if self.keep == "digest" else None
        if t["ev"] is not None:
            t["ev"][1].record(self.stream)
        done = torch.cuda.Event()
        done.record(self.stream)
        t["done"] = done
        self.inflight.append(t)
        return t

    def _complete(self, t: dict) -> None:
        """Waits for a launch and books its results, latencies and finished
        streams."""
        t["done"].synchronize()
        end = time.perf_counter()
        self.inflight.remove(t)
        items, owners, res, sizes = t["items"], t["owners"], t["res"], t["sizes"]
        if t["ev"] is not None:
            self.gpu_ms += t["ev"][0].elapsed_time(t["ev"][1])
        if items and self.trace is not None:
            from .trace import rows_from
            calls = np.concatenate([self.host_calls(key, lut) for key, lut in items])
            self.trace.calls(rows_from(calls.view(np.int32).reshape(-1, 8), res.cpu().numpy()))
        if items and self.keep:
            host = res.cpu().numpy() if self.keep is True else None
            dh = t["dig"].cpu().numpy() if t["dig"] is not None else None
            o = 0
            for k, ((s, pos), n) in enumerate(zip(owners, sizes)):
                if host is not None:
                    s.results.append(host[o:o + n].copy())
                else:
                    s.digests.append(int(dh[k]))
                    if pos in s.keep_at:
                        s.kept[pos] = res[o:o + n].cpu().numpy()
                o += n
        self.ticks += 1
        self.tick_ms.append((end - t["now"]) * 1e3)
        searched = {id(s): n for (s, _), n in zip(owners, sizes)}
        for s in t["streams"]:
            s.busy = False
            if id(s) in searched:
                s.latencies.append((end - s.ready_at) * 1e3)
                self.frames += 1
                self.calls += searched[id(s)]
            s.ready_at = end
            if s.done:
                s.release(self.eng)
                self.used -= stream_bytes(s.w, s.h)
                self.finished.append(s)
                self.active.remove(s)

    def tick(self, now: float) -> None:
        """One launch over the next picture of every active stream, waited for."""
        self._complete(self._launch(now, list(self.active)))

    def run(self, t0: float | None = None) -> float:
        """Serves until every submitted request is finished; arrival times are
        seconds after t0.  Returns the wall time used.

        Without pipelining each launch is waited for before the next is
        prepared.  With it (pipeline=True) two launches are in flight: while
        one runs on the GPU the host prepares and enqueues the next from the
        streams that are not in it (a stream's pictures stay one at a time:
        its next picture waits for its previous one), so host work overlaps
        device work and the streams fall into two alternating groups."""
        self.queue.sort(key=lambda s: s.arrival)
        t0 = time.perf_counter() if t0 is None else t0
        for s in self.queue:
            s.arrival += t0
        depth = 2 if self.pipeline else 1
        while self.queue or self.active or self.inflight:
            now = time.perf_counter()
            self._admit(now)
            ready = [s for s in self.active if not s.busy]
            if ready and len(self.inflight) < depth:
                self._launch(now, ready)
                continue
            if self.inflight:
                self._complete(self.inflight[0])
                continue
            if not self.queue:
                break
            wait = self.queue[0].arrival - now
            if wait > 0:
                time.sleep(min(wait, 0.05))
        return time.perf_counter() - t0

    def latencies(self) -> np.ndarray:
        return np.concatenate([np.array(s.latencies) for s in self.finished]) if self.finished \
            else np.zeros(0)


def poisson_requests(rate: float, duration: float, seed: int, clips: dict,
                     mix=(("240p", 0.30), ("720p", 0.40), ("1080p", 0.25), ("2160p", 0.05)),
                     frames=(30, 120), gop: str = "B") -> list[Stream]:
    """Requests of the serving config: Poisson arrivals, resolution mix,
    30..120 frames each (uniform), GOP4 B, sources from `clips`."""
    rng = np.random.default_rng(seed)
    out, t, sid = [], 0.0, 0
    names = [m[0] for m in mix]
    probs = np.array([m[1] for m in mix])
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            break
        name = names[int(rng.choice(len(names), p=probs / probs.sum()))]
        w, h = RES[name]
        nf = int(rng.integers(frames[0], frames[1] + 1))
        # the stream's pictures are consecutive clip pictures when the clip
        # is long enough (else it wraps round, which is a scene cut)
        span = 4 * max(1, (nf - 1 + 3) // 4) + 1 if gop == "B" else nf
        room = clips[name].frames - span
        off = int(rng.integers(0, room + 1)) if room >= 0 else int(rng.integers(0, 1000))
        out.append(Stream(sid, w, h, nf, gop, clips[name], offset=off, arrival=t))
        sid += 1
    return out


def run_alone(eng: me.Engine, s: Stream, keep="results"):
    """The same stream served by itself: its per-frame results (keep
    "results"), or its per-picture digests and the pictures of s.keep_at in
    full (keep "digest")."""
    solo = Stream(s.sid, s.w, s.h, s.nframes, s.gop, s.clip, s.offset,
                  keep_at=set(s.keep_at))
    srv = Server(eng, hbm_budget=1 << 62, keep_results=True if keep == "results" else "digest")
    srv.submit(solo)
    srv.run()
    return solo.results if keep == "results" else (solo.digests, solo.kept)
