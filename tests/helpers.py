"""Shared test inputs: golden traces and seeded frames."""
from __future__ import annotations

import glob
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def golden_traces():
    """[(name, refs{poc: plane}, curs{id: plane}, calls int32[N,16])]"""
    out = []
    for f in sorted(glob.glob(os.path.join(GOLDEN, "me_trace_*.npz"))):
        d = np.load(f)
        refs = {int(p): d["refs"][i] for i, p in enumerate(d["ref_pocs"])}
        curs = {int(p): d["curs"][i] for i, p in enumerate(d["cur_ids"])}
        name = os.path.basename(f)[len("me_trace_"):-4]
        out.append((name, refs, curs, d["calls"]))
    return out


def smooth_plane(h, w, seed, cell=8):
    """Bilinear-upsampled random lattice, float."""
    r = np.random.default_rng(seed)
    g = r.integers(0, 256, (h // cell + 2, w // cell + 2)).astype(float)
    yy = np.arange(h) / cell
    xx = np.arange(w) / cell
    y0 = yy.astype(int)
    x0 = xx.astype(int)
    fy = (yy - y0)[:, None]
    fx = (xx - x0)[None, :]
    a, b = g[y0][:, x0], g[y0][:, x0 + 1]
    c, d = g[y0 + 1][:, x0], g[y0 + 1][:, x0 + 1]
    # the fancy indexing above yields column-major arrays: hand back row-major
    return np.ascontiguousarray(a * (1 - fx) * (1 - fy) + b * fx * (1 - fy) +
                                c * (1 - fx) * fy + d * fx * fy)


def moving_pair(h, w, seed, shift=(3, -5), noise=3):
    """(ref, cur) uint8 planes: cur is ref's content shifted plus noise."""
    r = np.random.default_rng(seed + 7)
    s = smooth_plane(h, w, seed)
    ref = np.clip(s + r.integers(-noise, noise + 1, (h, w)), 0, 255).astype(np.uint8)
    cur = np.clip(np.roll(s, shift, (0, 1)) + r.integers(-noise, noise + 1, (h, w)), 0,
                  255).astype(np.uint8)
    return ref, cur


def random_calls(rng, w, h, count, sizes=(8, 16, 32, 64), srs=(16, 24, 32, 64, 4, 1),
                 slack=12):
    """int32 [count, 6] = px, py, cx, cy, n, sr with blocks inside the frame and
    centres anywhere within `slack` px outside it."""
    out = []
    sizes = [n for n in sizes if n <= w and n <= h]
    for _ in range(count):
        n = int(rng.choice(sizes))
        sr = int(rng.choice(srs))
        px = int(rng.integers(0, w - n + 1))
        py = int(rng.integers(0, h - n + 1))
        cx = int(rng.integers(-slack, w + slack))
        cy = int(rng.integers(-slack, h + slack))
        out.append((px, py, cx, cy, n, sr))
    return np.array(out, np.int32)
