// Host-only check of tcodec_amd/csrc/tcme_bounds.h at the int32 edges: the
// descriptor tests the kernels run must reject every out-of-range block
// without an intermediate sum overflowing.  Built with g++ -fsanitize=undefined
// by tests/test_bounds.py; prints "ok" and exits 0 when every check holds.
#include <climits>
#include <cstdio>

#include "tcme_bounds.h"

static int fails = 0;
#define CHECK(e)                                                        \
    do {                                                                \
        if (!(e)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #e); ++fails; } \
    } while (0)

int main()
{
    using namespace tcme;
    // in-frame blocks, edges included
    CHECK(block_fits(0, 0, 8, 8, 8));
    CHECK(block_fits(1912, 1072, 8, 1920, 1080));
    CHECK(!block_fits(1913, 0, 8, 1920, 1080));
    CHECK(!block_fits(-1, 0, 8, 1920, 1080));
    CHECK(!block_fits(0, 0, 16, 8, 64));          // block larger than the picture
    // the advisor's case: x near INT_MAX used to wrap x + n negative
    for (int x = INT_MAX - 70; x != INT_MIN; ++x) {
        CHECK(!block_fits(x, 0, 64, 1920, 1080));
        CHECK(!rect_fits(x, 0, 64, 1, 1920, 1080));
        if (x == INT_MAX) break;
    }
    CHECK(!block_fits(INT_MIN, INT_MIN, 8, 320, 240));
    CHECK(rect_fits(0, 0, 64, 64, 64, 64) && !rect_fits(1, 0, 64, 64, 64, 64));
    // spans of caller buffers
    CHECK(span_fits(0, 8, 8, 8, 64));
    CHECK(!span_fits(1, 8, 8, 8, 64));
    CHECK(!span_fits(LLONG_MAX - 10, 8, 8, 8, LLONG_MAX));
    CHECK(!span_fits(LLONG_MAX, 64, 64, INT_MAX, LLONG_MAX));
    CHECK(span_fits(0, 64, 64, INT_MAX, 64LL * INT_MAX));
    CHECK(!span_fits(-1, 8, 8, 8, 1000));
    CHECK(!span_fits(0, 8, 8, 7, 1000));           // stride below the width
    // the large-hexagon step: unchanged below the cap, capped above it, and
    // candidate coordinates bx + 4 * step stay inside int32
    CHECK(search_scale(24, 1920, 1080) == 6);
    CHECK(search_scale(3, 1920, 1080) == 1);
    CHECK(search_scale(-8, 1920, 1080) == 1);
    CHECK(search_scale(INT_MAX, 1920, 1080) == 1920);
    CHECK(search_scale(INT_MAX, 16384, 16384) == 16384);
    CHECK(16384LL + 4LL * search_scale(INT_MAX, 16384, 16384) <= INT_MAX);
    // with the capped step every large-hexagon point is off the picture, as
    // with the uncapped one
    {
        const int w = 320, h = 240, s = search_scale(INT_MAX, w, h);
        const int px[6] = {-2, 2, -4, 4, -2, 2}, py[6] = {-4, -4, 0, 0, 4, 4};
        for (int bx = 0; bx <= w - 8; bx += 13)
            for (int by = 0; by <= h - 8; by += 11)
                for (int k = 0; k < 6; ++k)
                    CHECK(!block_fits(bx + px[k] * s, by + py[k] * s, 8, w, h));
    }
    // centre tiles stay small for any raw centre
    CHECK(centre_tile(INT_MAX) == kMaxPicSide / 64);
    CHECK(centre_tile(INT_MIN) == 0);
    CHECK(centre_tile(130) == 2);
    CHECK(centre_tile(INT_MAX) * 64 + 64 + 40 > 0);
    if (fails) return 1;
    std::printf("ok\n");
    return 0;
}
