// Internal interface between the host runtime (tcme_api.hip) and the gfx950
// kernels (tcme_kernels.hip).  Not installed; the public ABI is include/tcme.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tcme.h"

namespace tcme {

// Device view of one picture.  The luma plane has `pitch` bytes per row and at
// least kTailSlack readable bytes after the last row, so a 4-byte read at any
// in-frame column may fetch the whole aligned 8-byte pair that covers it.
// `an` (reference pictures only) holds one 8-byte anchor per sample, `apitch`
// entries per row: bits 0-15 sample, 16-31 hh, 32-47 vv, 48-63 cc, with hh, vv, cc
// the raw (unclipped) six-tap half samples right of, below and diagonal to the
// sample.  Half samples whose taps would leave the frame hold 0 and are never
// read by the search or the predictor.
struct PicDev {
    const uint8_t *y;
    const uint64_t *an;
    int32_t w, h, pitch, apitch;
};

constexpr int kTailSlack = 256;
constexpr int kMaxBlk = 64;

// Launch shape of the search (chosen by tcme_autotune or set by the caller):
// calls per workgroup, wavefronts per workgroup, LDS window apron in pixels
// and LDS bytes per workgroup.  chunk 1 with no LDS is the plain form, one
// wavefront per call (me_search_kernel); it is the default.
struct Tiling {
    int chunk = 1;
    int waves = 1;
    int apron = 0;
    int lds_bytes = 0;
};

hipError_t launch_anchors(const PicDev &p, hipStream_t s);
// One plain-form kernel (one wavefront per call) for the calls of block size
// blk in the list; launch all four sizes to cover a mixed list.
hipError_t launch_search_n(int blk, const PicDev *table, int npics, const tcme_call *calls, int n,
                           tcme_result *res, tcme_stats *stats, hipStream_t s);
// The chunked kernel (Tiling::chunk > 1 or LDS in use).
hipError_t launch_search_tiled(const PicDev *table, int npics, const tcme_call *calls, int n,
                               const Tiling &t, tcme_result *res, tcme_stats *stats,
                               hipStream_t s);

}  // namespace tcme
