"""Seeded synthetic YCbCr 4:2:0 video for the motion-estimation workloads.

All randomness comes from a counter-based splitmix64 hash of (seed, stream,
index), so every host produces the same bytes for the same seed.

``gradient_video``  320x240-style clip: a diagonal luma gradient 0..255 plus
                    i.i.d. uniform integer noise in [-4, 4], the content
                    translated by (+3, -2) px per frame; chroma is 128 plus the
                    same noise sampled at the even luma positions.
``scene_video``     smooth value-noise background scaled to 16..235, K moving
                    textured rectangles (32..256 px per side, velocities in
                    [-16, 16] px/frame on the quarter-pel grid, wrapping at the
                    frame edge), optional global pan in px/frame (fractional
                    pans resample the background), Gaussian noise sigma 2.
"""
from __future__ import annotations

import numpy as np

_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_G = np.uint64(0x9E3779B97F4A7C15)


def mix64(x: np.ndarray) -> np.ndarray:
    """splitmix64 finaliser over a uint64 array (wrapping arithmetic)."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, np.uint64) + _G).astype(np.uint64)
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        return z ^ (z >> np.uint64(31))


def _key(seed: int, stream: int) -> np.uint64:
    return mix64(np.uint64((seed & 0xFFFFFFFF) << 20 | (stream & 0xFFFFF)))


def uniform01(seed: int, stream: int, idx: np.ndarray) -> np.ndarray:
    """Uniform [0,1) doubles, one per index."""
    k = _key(seed, stream)
    with np.errstate(over="ignore"):
        h = mix64(np.asarray(idx, np.uint64) ^ k)
    return (h >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))


def randint(seed: int, stream: int, idx: np.ndarray, lo: int, hi: int) -> np.ndarray:
    """Uniform integers in [lo, hi] (inclusive)."""
    k = _key(seed, stream)
    with np.errstate(over="ignore"):
        h = mix64(np.asarray(idx, np.uint64) ^ k)
    return (h % np.uint64(hi - lo + 1)).astype(np.int64) + lo


def _chroma_from(noise_or_luma: np.ndarray) -> np.ndarray:
    return noise_or_luma[::2, ::2]


def gradient_video(w: int = 320, h: int = 240, frames: int = 8, seed: int = 1,
                   shift: tuple[int, int] = (3, -2), noise: int = 4):
    """List of (Y, Cb, Cr) uint8 planes."""
    yy, xx = np.mgrid[0:h, 0:w].astype(np.int64)
    out = []
    for t in range(frames):
        sx, sy = shift[0] * t, shift[1] * t
        g = ((xx - sx) + (yy - sy)) * 255 // (w + h - 2)
        idx = (yy * w + xx).astype(np.uint64) + np.uint64(t) * np.uint64(1 << 32)
        nz = randint(seed, 1, idx, -noise, noise)
        y = np.clip(g + nz, 0, 255).astype(np.uint8)
        c = np.clip(128 + _chroma_from(nz), 0, 255).astype(np.uint8)
        out.append((y, c.copy(), c.copy()))
    return out


def _value_noise(seed: int, w: int, h: int, ox: float, oy: float,
                 periods=(128, 48, 16), amps=(1.0, 0.5, 0.25)) -> np.ndarray:
    """Smooth fractal value noise sampled at (x + ox, y + oy), range ~[0, 1]."""
    xs = np.arange(w, dtype=np.float64) + ox
    ys = np.arange(h, dtype=np.float64) + oy
    acc = np.zeros((h, w))
    for o, (p, a) in enumerate(zip(periods, amps)):
        gx = np.floor(xs / p).astype(np.int64)
        gy = np.floor(ys / p).astype(np.int64)
        fx = xs / p - gx
        fy = ys / p - gy
        fx = fx * fx * (3 - 2 * fx)
        fy = fy * fy * (3 - 2 * fy)
        x0, x1 = gx.min(), gx.max() + 1
        y0, y1 = gy.min(), gy.max() + 1
        ly, lx = np.mgrid[y0:y1 + 1, x0:x1 + 1]
        lid = ((ly + (1 << 20)).astype(np.uint64) << np.uint64(24)) ^ (lx + (1 << 20)).astype(np.uint64)
        lat = uniform01(seed, 100 + o, lid)
        ix = gx - x0
        iy = gy - y0
        v00 = lat[iy][:, ix]
        v10 = lat[iy][:, ix + 1]
        v01 = lat[iy + 1][:, ix]
        v11 = lat[iy + 1][:, ix + 1]
        top = v00 + (v10 - v00) * fx[None, :]
        bot = v01 + (v11 - v01) * fx[None, :]
        acc += a * (top + (bot - top) * fy[:, None])
    return acc / sum(amps)


class Scene:
    """Moving-rectangle scene; frame(t) renders display frame t."""

    def __init__(self, w: int, h: int, seed: int, rects: int = 16,
                 pan: tuple[float, float] = (0.0, 0.0), sigma: float = 2.0):
        self.w, self.h, self.seed, self.pan, self.sigma = w, h, seed, pan, sigma
        i = np.arange(rects, dtype=np.uint64)
        self.rw = randint(seed, 10, i, 32, 256)
        self.rh = randint(seed, 11, i, 32, 256)
        self.x0 = uniform01(seed, 12, i) * (w + self.rw) - self.rw
        self.y0 = uniform01(seed, 13, i) * (h + self.rh) - self.rh
        self.vx = randint(seed, 14, i, -64, 64) / 4.0
        self.vy = randint(seed, 15, i, -64, 64) / 4.0
        self.base = randint(seed, 16, i, 40, 215).astype(np.float64)
        self.tex = []
        for k in range(rects):
            tw, th = int(self.rw[k]) // 8 + 2, int(self.rh[k]) // 8 + 2
            tid = np.arange(tw * th, dtype=np.uint64) + np.uint64(k << 24)
            self.tex.append((uniform01(seed, 17, tid).reshape(th, tw) * 2 - 1) * 40.0)

    def luma_float(self, t: int) -> np.ndarray:
        w, h = self.w, self.h
        bg = _value_noise(self.seed, w, h, -self.pan[0] * t, -self.pan[1] * t)
        img = 16.0 + bg * (235.0 - 16.0)
        for k in range(len(self.tex)):
            rw, rh = int(self.rw[k]), int(self.rh[k])
            px = (self.x0[k] + self.vx[k] * t + rw) % (w + rw) - rw
            py = (self.y0[k] + self.vy[k] * t + rh) % (h + rh) - rh
            xa, xb = max(0, int(np.ceil(px))), min(w, int(np.ceil(px + rw)))
            ya, yb = max(0, int(np.ceil(py))), min(h, int(np.ceil(py + rh)))
            if xa >= xb or ya >= yb:
                continue
            u = (np.arange(xa, xb) - px) / 8.0
            v = (np.arange(ya, yb) - py) / 8.0
            iu = np.minimum(np.floor(u).astype(np.int64), self.tex[k].shape[1] - 2)
            iv = np.minimum(np.floor(v).astype(np.int64), self.tex[k].shape[0] - 2)
            fu = (u - iu)[None, :]
            fv = (v - iv)[:, None]
            T = self.tex[k]
            t00 = T[iv][:, iu]
            t10 = T[iv][:, iu + 1]
            t01 = T[iv + 1][:, iu]
            t11 = T[iv + 1][:, iu + 1]
            top = t00 + (t10 - t00) * fu
            bot = t01 + (t11 - t01) * fu
            img[ya:yb, xa:xb] = self.base[k] + top + (bot - top) * fv
        return img

    def frame(self, t: int):
        lum = self.luma_float(t)
        h, w = self.h, self.w
        idx = np.arange(h * w, dtype=np.uint64).reshape(h, w) + np.uint64(t) * np.uint64(1 << 34)
        u1 = np.maximum(uniform01(self.seed, 20, idx), 1e-300)
        u2 = uniform01(self.seed, 21, idx)
        g = np.sqrt(-2.0 * np.log(u1)) * np.cos(2 * np.pi * u2) * self.sigma
        y = np.clip(np.rint(lum + g), 0, 255).astype(np.uint8)
        c = np.clip(np.rint(128 + (lum[::2, ::2] - 128) / 8.0), 0, 255).astype(np.uint8)
        return y, c, c.copy()


def scene_video(w: int, h: int, frames: int, seed: int, rects: int = 16,
                pan: tuple[float, float] = (0.0, 0.0)):
    s = Scene(w, h, seed, rects, pan)
    return [s.frame(t) for t in range(frames)]


def write_yuv(path: str, frames) -> None:
    with open(path, "wb") as f:
        for y, cb, cr in frames:
            f.write(np.ascontiguousarray(y).tobytes())
            f.write(np.ascontiguousarray(cb).tobytes())
            f.write(np.ascontiguousarray(cr).tobytes())
