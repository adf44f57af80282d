/*
 * tcme.h - MI355X (gfx950) motion-estimation backend for TCodec.
 *
 * Two faces, one library (libtcme.so):
 *
 * 1. The drop-in.  tc_motion_est() has exactly the reference's signature and
 *    result (include/tcodec_common.h:454-459, src/motion.c:384-485 in the
 *    reference tree).  Link it in place of motion.c's own definition and the
 *    encoder's .tcv output is unchanged.  It never falls back to the CPU: if
 *    no GPU can be used it prints the reason and aborts.
 *
 * 2. The batch API (tcme_*).  Pictures live in device memory; a call list
 *    (any mix of references, current pictures, block sizes and search ranges)
 *    is searched in one launch on a caller-supplied HIP stream.  Call lists,
 *    results and statistics are device pointers.  No function here allocates
 *    inside tcme_search(); pictures are allocated by tcme_pic_create().
 *
 * All entry points are C ABI with plain pointers and sizes.  Errors are
 * negative return codes; tcme_last_error() gives the message (thread-local).
 */
#ifndef TCME_H
#define TCME_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- drop-in boundary ------------------------------------------------- */

/* Same layout as the reference's tc_mv_s (include/tcodec_types.h:168-171). */
typedef struct tcme_mv { int32_t x, y; } tcme_mv;

/* Replaces tc_motion_est, reference src/motion.c:384-389 (prototype
 * include/tcodec_common.h:454-459).  Returns the absolute quarter-pel
 * position of the best n x n block and, if best_sad_out is not NULL, its SAD.
 * Any blk_size >= 1, as the reference (8, 16, 32 and 64, the sizes its
 * encoders use, run on kernels specialised for them; other sizes on a
 * generic kernel).
 * A translation unit that also includes the reference's tcodec_common.h,
 * which declares the same symbol with its own (layout-identical) types,
 * defines TCME_HIDE_DROPIN before including this header. */
#ifndef TCME_HIDE_DROPIN
tcme_mv tc_motion_est(const uint8_t *ref, int ref_stride, int ref_w, int ref_h,
                      const uint8_t *orig, int orig_stride, int center_x, int center_y,
                      int blk_size, int search_range, int32_t *best_sad_out);
#endif

/* ---- batch API ---------------------------------------------------------- */

typedef struct tcme_ctx tcme_ctx;

/* One search: reference picture, current picture, block origin (px,py) in
 * the current picture, full-pel centre, block size, search range.  Mirrors
 * the arguments of the reference call with the two planes named by id. */
typedef struct tcme_call {
    int32_t ref_pic, cur_pic;
    int32_t px, py;
    int32_t cx, cy;
    int32_t blk, sr;
} tcme_call;

typedef struct tcme_result {
    int32_t mv_x, mv_y;   /* absolute quarter-pel position */
    int32_t sad;
    int32_t status;       /* 0 ok, <0 rejected descriptor (see TCME_E_*) */
} tcme_result;

/* Per-call candidate counts in the reference's sequential order, used for
 * the algorithmic-byte figures: in-frame integer SADs, 6-tap quarter-pel
 * SADs, bilinear quarter-pel SADs, out-of-frame integer candidates. */
typedef struct tcme_stats {
    int32_t int_evals, q6_evals, qbil_evals, int_oob;
} tcme_stats;

enum {
    TCME_OK = 0,
    TCME_E_ARG = -1,
    TCME_E_HIP = -2,
    TCME_E_NOMEM = -3,
    TCME_E_PIC = -4,
    TCME_E_BLK = -5
};

const char *tcme_last_error(void);
const char *tcme_version(void);

/* Context on a device (-1 = current).  max_pics bounds the picture table. */
int  tcme_create(tcme_ctx **out, int device, int max_pics);
void tcme_destroy(tcme_ctx *ctx);

/* Allocates a w x h luma picture with its anchor plane (per sample: the sample
 * and its raw six-tap half samples right, below and diagonal) on the device. */
int  tcme_pic_create(tcme_ctx *ctx, int w, int h);
/* A plain w x h plane with no anchor plane: current pictures and chroma
 * planes (it cannot be a search or luma-prediction reference). */
int  tcme_pic_create_plane(tcme_ctx *ctx, int w, int h);
int  tcme_pic_destroy(tcme_ctx *ctx, int pic);
/* Fills a picture from host (upload) or device memory, then rebuilds its
 * anchor plane (if it has one).  Asynchronous on stream (a hipStream_t); an
 * upload has copied the host plane into pinned staging by the time it returns,
 * so the caller may reuse that buffer at once. */
int  tcme_pic_upload(tcme_ctx *ctx, int pic, const uint8_t *host, int stride, void *stream);
int  tcme_pic_upload_device(tcme_ctx *ctx, int pic, const uint8_t *dev, int stride,
                            void *stream);
/* Device address and pitch of a picture's luma plane (for callers that
 * write pixels themselves; call tcme_pic_finish afterwards). */
int  tcme_pic_luma(tcme_ctx *ctx, int pic, uint8_t **dev, int *pitch);
int  tcme_pic_finish(tcme_ctx *ctx, int pic, void *stream);

/* Searches n calls (device array) into results (device array); stats may be
 * NULL.  Asynchronous on stream.  Returns TCME_OK once launched. */
int  tcme_search(tcme_ctx *ctx, const tcme_call *calls, int n, tcme_result *results,
                 tcme_stats *stats, void *stream);

/* Launch shape of tcme_search.  Results never depend on it.
 *   waves8..waves64  calls searched one per workgroup: wavefronts per call
 *                    by block size (8x8: 1; 16x16: 1; 32x32: 1, 2 or 4;
 *                    64x64: 2, 4 or 8)
 *   tile_waves       4 or 8: the 8x8 and 16x16 calls are searched by tile
 *                    workgroups of that many wavefronts instead; each takes
 *                    tile_chunk (1..96) consecutive calls, stages in LDS the
 *                    reference around each run of calls that share a
 *                    reference picture and a 64x64 centre tile, and searches
 *                    many calls per wavefront from there.  Lists grouped by
 *                    CTU and reference (85 calls per group) suit chunk 85.
 *                    0: no tile workgroups.
 *   tile_big         1 (with tile_waves 4 or 8): the tile workgroups take
 *                    every call, all block sizes (one launch)
 * A new context starts with {1, 1, 1, 4, 8, 85, 1}: every call in one launch
 * of 8-wavefront tile workgroups, 85 calls each. */
typedef struct tcme_tiling {
    int32_t waves8, waves16, waves32, waves64;
    int32_t tile_waves, tile_chunk;
    int32_t tile_big;
} tcme_tiling;

int  tcme_set_tiling(tcme_ctx *ctx, const tcme_tiling *t);
int  tcme_get_tiling(tcme_ctx *ctx, tcme_tiling *t);

/* Times a grid of launch shapes on the caller's call list (device arrays;
 * results are overwritten with the same answers) and keeps the fastest in
 * the context.  Synchronises the stream. */
int  tcme_autotune(tcme_ctx *ctx, const tcme_call *calls, int n, tcme_result *results,
                   int reps, void *stream, tcme_tiling *best, float *best_ms);

/* A search captured once as a HIP graph and replayed per frame: the launches
 * tcme_search makes for (calls, n, results, stats) with the context's launch
 * shape at capture time, replayed with no per-launch host work.  The buffers
 * are read when a replay runs, so the caller rewrites the call list (same n)
 * and the pictures between replays.  Replays on one graph must not overlap,
 * and per-kernel timing (tcme_set_timing) does not see them.  The graph
 * belongs to the context: destroy it first. */
typedef struct tcme_graph tcme_graph;
int  tcme_graph_create(tcme_ctx *ctx, const tcme_call *calls, int n, tcme_result *results,
                       tcme_stats *stats, tcme_graph **out);
int  tcme_graph_launch(tcme_graph *g, void *stream);
void tcme_graph_destroy(tcme_graph *g);

/* ---- ME-only predictor chain ------------------------------------------- */

/* The reference encoder centres every search at the min-magnitude predictor
 * qt_mvp (src/encoder.c:405-427) of a per-CTU 8x8 grid (reset per CTU,
 * :1425-1428) that its RDO fills with each leaf's final displacement
 * (:1279-1288), visiting nodes in qt_split's decision order (:1292-1397).
 * The chain runs that dependency on the device without the RDO: one
 * workgroup per CTU walks the nodes in that order, centres each leaf's
 * searches (one per reference, in order) at qt_mvp of the grid, keeps the
 * cheapest J = SAD + lambda * (se_bits(mvd.x) + se_bits(mvd.y)) (first
 * reference on ties; se_bits as src/bitstream.c:100-112; mvd = mv - (mvp +
 * 4 pos)), fills the leaf's cells with the chosen displacement, and keeps the
 * four children when their J sum is not above the parent's.  Every search is
 * tc_motion_est's own; the decisions are the chain's (ME-only), so the calls
 * are those an encoder with this decision rule would make. */
typedef struct tcme_chain_frame {
    int32_t cur_pic, nrefs;   /* 1..4 references, searched in this order */
    int32_t ref_pic[4];
    int32_t sr, lambda;       /* search range; lambda >= 0 */
} tcme_chain_frame;

/* One search of the chain: reference index, node, centre, block size,
 * result. */
typedef struct tcme_chain_call {
    int32_t ref, px, py, cx, cy, blk;
    int32_t mv_x, mv_y, sad, reserved;
} tcme_chain_call;

/* Records per CTU in the output: at most 169 leaves (1 + 4 + 4 x (1 + 4 + 4 x (1 + 4 + 4))) x 4 references. */
#define TCME_CHAIN_MAX_CALLS 676

/* Runs the chain over n frames (host array, 1 <= n <= 16; every reference the
 * size of its current picture).  The frames' CTUs are numbered frame after
 * frame, raster order, ceil(w/64) x ceil(h/64) per frame; CTU k's searches go
 * to calls[k * TCME_CHAIN_MAX_CALLS ...] in the order they are made and their
 * number to counts[k].  calls and counts are device arrays with room for
 * ctu_cap CTUs.  Asynchronous on stream. */
int  tcme_search_chain(tcme_ctx *ctx, const tcme_chain_frame *frames, int n,
                       tcme_chain_call *calls, int32_t *counts, int64_t ctu_cap, void *stream);

/* ---- motion compensation (batched) -------------------------------------- */

/* One predicted block: picture id, absolute motion vector (quarter sample
 * for luma, eighth sample on a chroma plane), block size (1..64), and where
 * the block goes in the caller's output buffer (byte offset of its first
 * sample, row stride). */
typedef struct tcme_pred {
    int32_t pic, mv_x, mv_y, blk;
    int64_t out_off;
    int32_t out_stride, reserved;
} tcme_pred;

/* Replaces tc_inter_predict (reference src/motion.c:493-584, prototype
 * include/tcodec_common.h:464-468) for n blocks at once: pic must be a
 * reference picture (tcme_pic_create).  preds, out and status (optional, one
 * TCME_* code per block) are device arrays; out holds out_bytes bytes.
 * Asynchronous on stream. */
int  tcme_predict_luma(tcme_ctx *ctx, const tcme_pred *preds, int n, uint8_t *out,
                       int64_t out_bytes, int32_t *status, void *stream);
/* Replaces tc_inter_predict_chroma (reference src/motion.c:597-632,
 * include/tcodec_common.h:476-480) for n blocks: pic is a chroma plane
 * (tcme_pic_create_plane of the chroma size). */
int  tcme_predict_chroma(tcme_ctx *ctx, const tcme_pred *preds, int n, uint8_t *out,
                         int64_t out_bytes, int32_t *status, void *stream);

/* ---- encoder-side helpers (batched) ------------------------------------ */

/* Replaces estimate_global_mv (reference src/encoder.c:171-202) for n
 * (current, reference) picture pairs of one size: pairs holds 2 ids per pair,
 * out receives 3 ints per pair (glob_mv_x, glob_mv_y in quarter-pel, status).
 * Device arrays, asynchronous on stream; n <= 65535.  The context keeps a
 * workspace for the coarse SADs of every displacement within +-128 px (grown
 * on first use, freed with the context); successive calls on one context are
 * ordered on the device whatever their streams. */
int  tcme_global_mv(tcme_ctx *ctx, const int32_t *pairs, int n, int32_t *out, void *stream);

/* tcme_global_mv with no allocation at all: no counterpart in the reference;
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The coarse
 * SADs go to a fixed 16 MiB scratch array in device memory (one per GPU, held
 * until the process ends) instead of the context's workspace; a call whose
 * table would not fit (n x bands x 4,225 x 4 bytes, the bands set by the
 * context's largest picture: about 152 KB a pair at 1920x1080, 574 KB at
 * 3840x2160) fails with TCME_E_ARG.  Calls on one context are ordered on the
 * device; calls on different contexts of one GPU that overlap in time race
 * on the shared array. */
int  tcme_global_mv_static(tcme_ctx *ctx, const int32_t *pairs, int n, int32_t *out,
                           void *stream);

/* A block SAD: the w x h block (1..64 each) of picture pic at (x,y) against
 * the block of picture other_pic at (ox,oy), or, with other_pic < 0, against
 * the caller's buffer at buf_off with row stride buf_stride.  Covers the SAD
 * proxies of the reference encoder (tc_sad of a prediction,
 * src/encoder.c:814, 839, 955, 1171) and its co-located stability SADs
 * (src/encoder.c:2843-2891). */
typedef struct tcme_sad_desc {
    int32_t pic, x, y, w, h;
    int32_t other_pic, ox, oy;
    int64_t buf_off;
    int32_t buf_stride, reserved;
} tcme_sad_desc;

/* n SADs into out (>= 0, or a negative TCME_* code for a bad descriptor);
 * descs, buf (may be NULL when no descriptor uses it, buf_bytes long) and
 * out are device arrays.  Asynchronous on stream. */
int  tcme_sad(tcme_ctx *ctx, const tcme_sad_desc *descs, int n, const uint8_t *buf,
              int64_t buf_bytes, int32_t *out, void *stream);

/* Per-kernel timing of dispatches (events on the stream each kernel runs
 * on).  tcme_last_timing waits for the last dispatch and writes the duration
 * in ms of each kernel it launched, in launch order: the 64x64 and 32x32
 * per-call kernels, then the tile kernel (or the 16x16 and 8x8 per-call
 * kernels); a single tile kernel with tile_big.  Returns the kernel count. */
int  tcme_set_timing(tcme_ctx *ctx, int on);
int  tcme_last_timing(tcme_ctx *ctx, float *ms, int max);

/* ---- batched time steps ------------------------------------------------ */

/* One picture of a time step: its current picture, references (searched in
 * this order) and search range. */
typedef struct tcme_tick_pic {
    int32_t cur_pic, nrefs, sr;
    int32_t ref_pic[4];
} tcme_tick_pic;

/* The batched and served forms' time step from C: every in-frame quadtree
 * node (64, 32, 16, 8) of each picture against each of its references,
 * centred at the block position, searched in ONE launch.  The call list is
 * built on the device from per-shape layouts the context keeps (built once
 * per picture size, reference count and range); pictures' calls follow one
 * another, offsets (host, n + 1 entries) receive each picture's first result
 * index and the total; results (device) holds results_cap entries.  pics is
 * a host array.  Asynchronous on stream; a context's steps share one call
 * list, so successive steps go on one stream (or the caller orders them). */
int  tcme_tick_search(tcme_ctx *ctx, const tcme_tick_pic *pics, int n, tcme_result *results,
                      int64_t results_cap, int64_t *offsets, void *stream);
/* A picture shape's call list as tcme_tick_search lays it out (complete CTUs
 * first, then partial ones at the edges; reference-major; CTU raster order;
 * largest blocks first inside a CTU, each size in raster order), with
 * ref_pic = the reference index and cur_pic = -1.  Returns the call count;
 * calls (host, cap entries) may be NULL, and so may ctx (no device work). */
int64_t tcme_tick_layout(tcme_ctx *ctx, int w, int h, int nrefs, int sr, tcme_call *calls,
                         int64_t cap);

/* Host-pointer convenience: copies calls in, searches, copies results out,
 * synchronises the stream. */
int  tcme_search_host(tcme_ctx *ctx, const tcme_call *calls, int n, tcme_result *results,
                      tcme_stats *stats, void *stream);

/* ---- frame-level drop-in ------------------------------------------------ */

/* The reference encoder calls tc_motion_est (src/encoder.c:807, 834, 928,
 * 982, 1004, 1035) once per quadtree node and reference, sequentially from
 * its RDO, on planes that do not change during a frame (the current picture
 * is staged at src/encoder.c:2631, the picture buffer is updated only after
 * the frame, src/encoder.c:3085-3097).  The frame-level form replaces those
 * calls (same arguments, same results) with:
 *   tcme_frame_begin        once per coded picture: its planes and the
 *                           searches whose centres are known at frame start,
 *                           which are all searched at once, in one launch;
 *   tcme_frame_motion_est   every call: answered from those results when it
 *                           is one of them; otherwise searched at once on the
 *                           resident planes, in one launch with the other
 *                           nodes of its 64x64 CTU against the same
 *                           references at the same displacement (the
 *                           encoder's predictor, src/encoder.c:405-427, is
 *                           reset per CTU, src/encoder.c:1425-1428, so these
 *                           are its likely next calls).
 * It has no counterpart in the reference: include the hook file
 * tcodec_amd/glue/tcme_encoder_hook.c to drive it from the unmodified
 * encoder.  Every answer equals tc_motion_est's. */
typedef struct tcme_frame tcme_frame;

/* Speculative centres of one set of searches. */
enum { TCME_CENTRE_OFFSET = 0, TCME_CENTRE_GRID = 1 };

/* One 8x8 cell of a displacement grid (quarter samples), row-major. */
typedef struct tcme_cell {
    int16_t dx, dy;
    uint8_t unavailable;  /* nonzero: no search for nodes whose top-left cell this is */
    uint8_t reserved;
} tcme_cell;

/* Every in-frame quadtree node of the sizes in `sizes` (bit k: 8 << k)
 * against reference `ref` with search range `sr`, centred as the reference
 * encoder centres a call on a quarter-sample displacement d: the node at
 * (px, py) gets the full-sample centre ((4 px + d.x) >> 2, (4 py + d.y) >> 2).
 * OFFSET: d = (dx, dy) for every node (0: the block position, as at the
 * predictor sites when the predictor is zero; the global-motion estimate, as
 * at src/encoder.c:1000-1004).  GRID: d = the cell of grid_w x grid_h `grid`
 * at (px/8, py/8) (the previous frame's displacements, as at
 * src/encoder.c:1023-1035); nodes outside the grid or on an unavailable cell
 * get no search.  The grid is read during tcme_frame_begin only. */
typedef struct tcme_spec {
    int32_t ref, sr, sizes, mode;
    int32_t dx, dy;
    const tcme_cell *grid;
    int32_t grid_w, grid_h;
} tcme_spec;

/* Counters of a frame object since its creation. */
typedef struct tcme_frame_stats {
    int64_t frames;       /* tcme_frame_begin calls */
    int64_t spec_calls;   /* searches launched at frame start */
    int64_t calls;        /* tcme_frame_motion_est calls */
    int64_t hits_spec;    /* answered from the frame-start searches */
    int64_t hits_batch;   /* answered from an earlier miss's launch */
    int64_t misses;       /* searched at once (one launch each) */
    int64_t batch_calls;  /* searches those launches made */
    int64_t uncached;     /* not on the declared planes: handed to tc_motion_est */
    double begin_s;       /* host seconds inside tcme_frame_begin */
    double wait_s;        /* host seconds waiting for the frame-start results */
    double miss_s;        /* host seconds inside misses */
} tcme_frame_stats;

/* Starts the GPU runtime and the drop-in's device context on a background
 * thread and returns at once, so that an encoder can overlap the runtime's
 * start with its first (intra) picture.  Later calls do nothing. */
void tcme_prepare(void);

int  tcme_frame_create(tcme_frame **out);
void tcme_frame_destroy(tcme_frame *f);
/* Declares the planes of the picture about to be coded (host memory, w x h
 * luma; refs[i] may be NULL) and launches the searches `specs` describe.  The
 * planes must not change until the next tcme_frame_begin on this object.
 * Returns once the work is queued. */
int  tcme_frame_begin(tcme_frame *f, const uint8_t *cur, int cur_stride, int w, int h,
                      const uint8_t *const *refs, const int *ref_strides, int nrefs,
                      const tcme_spec *specs, int nspecs);
/* Ends the declared frame (its planes may change from now on): until the
 * next tcme_frame_begin every call is handed to tc_motion_est. */
int  tcme_frame_end(tcme_frame *f);
/* tc_motion_est for the declared frame: the arguments and result of
 * reference src/motion.c:384-389.  A call that is not a node of the declared
 * planes (another reference, a block outside the current picture such as the
 * v1 coder's stack copy, an unaligned position) is handed to tc_motion_est.
 * Thread-safe (calls on one object are serialised). */
tcme_mv tcme_frame_motion_est(tcme_frame *f, const uint8_t *ref, int ref_stride, int ref_w,
                              int ref_h, const uint8_t *orig, int orig_stride, int center_x,
                              int center_y, int blk_size, int search_range,
                              int32_t *best_sad_out);
int  tcme_frame_get_stats(tcme_frame *f, tcme_frame_stats *st);

#ifdef __cplusplus
}
#endif
#endif
