// Descriptor bounds of the search, prediction and SAD kernels, written so that
// no sum can overflow whatever int32 values a caller passes: every test is of
// the form `x <= limit - size` with both sides in range, and sizes are checked
// before they are subtracted.  Plain integer code, usable from the kernels
// (hipcc) and from a host-only unit test (g++).
#pragma once
#include <stdint.h>

#if defined(__HIPCC__)
#define TCME_HD __host__ __device__ __forceinline__
#else
#define TCME_HD static inline
#endif

namespace tcme {

constexpr int kMaxPicSide = 16384;

// An n x n block at (x, y) lies inside a w x h picture (n >= 1).
TCME_HD bool block_fits(int x, int y, int n, int w, int h)
{
    return x >= 0 && y >= 0 && n <= w && n <= h && x <= w - n && y <= h - n;
}

// A bw x bh rectangle at (x, y) lies inside a w x h picture (bw, bh >= 1).
TCME_HD bool rect_fits(int x, int y, int bw, int bh, int w, int h)
{
    return x >= 0 && y >= 0 && bw <= w && bh <= h && x <= w - bw && y <= h - bh;
}

// Rows of bw bytes, bh rows `stride` apart from byte `off`, lie inside a
// buffer of `bytes` bytes (1 <= bw, bh <= 64; stride >= bw checked first).
TCME_HD bool span_fits(int64_t off, int bw, int bh, int stride, int64_t bytes)
{
    if (stride < bw || off < 0 || bytes < 0) return false;
    const int64_t span = static_cast<int64_t>(bh - 1) * stride + bw;
    return span <= bytes && off <= bytes - span;
}

// Large-hexagon step of a search (reference src/motion.c:398: sr/4, at least
// 1), capped at the picture's larger side.  The cap changes nothing: with a
// step of max(w, h) or more every large-hexagon point, (+-2, +-4) and (+-4, 0)
// steps away, lies off the picture, exactly as with the uncapped step, so the
// walk, its result and its candidate counts are the same; it keeps the
// candidate coordinates inside int32.
TCME_HD int search_scale(int sr, int w, int h)
{
    int s = sr / 4 > 1 ? sr / 4 : 1;
    const int m = w > h ? w : h;
    const int cap = m > 1 ? m : 1;
    return s > cap ? cap : s;
}

// The 64x64 tile column (or row) of a raw centre coordinate.  The coordinate
// is clamped into [0, kMaxPicSide] first (the search clamps centres into the
// picture anyway), so tile * 64 + 64 + a margin stays inside int32.
TCME_HD int centre_tile(int c)
{
    c = c < 0 ? 0 : (c > kMaxPicSide ? kMaxPicSide : c);
    return c >> 6;
}

}  // namespace tcme
