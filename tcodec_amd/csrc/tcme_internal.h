// Internal interface between the host runtime (tcme_api.hip) and the gfx950
// kernels (tcme_kernels.hip).  Not installed; the public ABI is include/tcme.h.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "tcme.h"
#include "tcme_bounds.h"

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
constexpr int kTileChunkMax = 96;

// Launch shape of the search (chosen by tcme_autotune or set by the caller):
// wavefronts per call for each block size (8x8 and 16x16: 1; 32x32: 1, 2
// or 4; 64x64: 2, 4 or 8), and the tile workgroups that take the 8x8 and
// 16x16 calls instead when tile_waves is 4 or 8 (and every call with
// tile_big).
struct Tiling {
    int waves[4] = {1, 1, 1, 4};  // for block sizes 8, 16, 32, 64
    int tile_waves = 8;
    int tile_chunk = 85;
    int tile_big = 1;             // tile workgroups also take 32x32/64x64: one launch
};

bool valid_waves(int blk, int waves);
// Compute units of the current device (cached per device; 256 if unknown).
int device_cus();
// Global-motion estimates and block SADs (tcme_aux.hip).  The estimate
// needs gmv_part_bytes(n, W, H) bytes of workspace, W x H bounding every
// picture a pair may name.
size_t gmv_part_bytes(int n, int max_w, int max_h);
hipError_t launch_global_mv(const PicDev *table, int npics, const int32_t *pairs, int n,
                            int max_w, int max_h, int32_t *part, int32_t *out, hipStream_t s);
hipError_t launch_sad(const PicDev *table, int npics, const tcme_sad_desc *descs, int n,
                      const uint8_t *buf, int64_t buf_bytes, int32_t *out, hipStream_t s);
// Motion compensation of n blocks (tcme_mc.hip): luma from the anchor plane,
// or chroma from a plain plane.
hipError_t launch_predict(bool chroma, const PicDev *table, int npics, const tcme_pred *preds,
                          int n, uint8_t *out, int64_t out_bytes, int32_t *status, hipStream_t s);
// The tile kernel: tile workgroups of `waves` wavefronts, `chunk` calls each;
// searches the calls of size 8 and 16 in the list, or, with `big` (4
// wavefronts), every call.
hipError_t launch_tile(int waves, bool big, int chunk, const PicDev *table, int npics,
                       const tcme_call *calls, int n, tcme_result *res, tcme_stats *stats,
                       hipStream_t s);
hipError_t launch_anchors(const PicDev &p, hipStream_t s);
// The search kernel for the calls of block size blk in the list (the others
// are skipped); launch all four sizes to cover a mixed list.
// One drop-in call (c.blk 8, 16, 32 or 64; c.ref_pic a reference picture of
// `table`): `blk` and `res` are device-visible pointers, normally to mapped
// pinned host memory; the current block is c.blk x c.blk bytes of pitch c.blk,
// followed by at least kTailSlack readable bytes.
hipError_t launch_dropin(const PicDev *table, const tcme_call &c, const uint8_t *blk,
                         tcme_result *res, hipStream_t s);
// A picture of a tick (tcme_tick_search): its layout (device template whose
// ref_pic fields are reference indices), picture ids and place in the list.
struct TickPicDev {
    const tcme_call *tmpl;
    int64_t first;
    int32_t ncalls, cur_pic;
    int32_t ref_pic[4];
};
hipError_t launch_expand(const TickPicDev *pics, int npics, tcme_call *out, int64_t total,
                         hipStream_t s);

// The ME-only predictor chain: up to kChainFrames frames per launch, one
// workgroup per CTU (frames' CTUs numbered from `first`, raster order, `cols`
// CTUs per row), kChainMaxCalls records per CTU in `calls`.  Jobs are
// validated by the host.
constexpr int kChainFrames = 16;
constexpr int kChainMaxCalls = 676;
struct ChainJobHost {
    int cur_pic, nrefs, ref_pic[4], sr, lambda, first, cols;
};
hipError_t launch_chain(const PicDev *table, const ChainJobHost *jobs, int njobs, int total_ctus,
                        tcme_chain_call *calls, int32_t *counts, hipStream_t s);

// n validated calls on resident pictures (calls and res device-visible,
// normally mapped host memory), one workgroup per call, any supported size;
// each result's status is written last, after a system-scope fence.
hipError_t launch_list(const PicDev *table, const tcme_call *calls, int n, tcme_result *res,
                       hipStream_t s);
hipError_t launch_search_n(int blk, int waves, const PicDev *table, int npics,
                           const tcme_call *calls, int n, tcme_result *res, tcme_stats *stats,
                           hipStream_t s);

}  // namespace tcme
