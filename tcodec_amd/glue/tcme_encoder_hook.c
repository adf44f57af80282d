/*
 * tcme_encoder_hook.c - drives libtcme.so's frame-level drop-in from the
 * reference encoder without editing it.
 *
 * Build: compile the reference's src/encoder.c with two
 * renames on its command line,
 *     -Dtc_ratectl_frame_start=tcme_hook_frame_start
 *     -Dtc_motion_est=tcme_hook_motion_est
 * compile src/motion.c with -Dtc_motion_est=tc_motion_est_ref (its CPU search
 * is then unused and the library's tc_motion_est serves the rest), compile
 * this file with the reference's include/ on the include path, and link
 * everything with -ltcme.
 *
 * tcme_hook_frame_start runs the real tc_ratectl_frame_start
 * (src/ratectl.c:89), which encode_poc_frame calls once per coded picture with
 * &enc->rc after staging the source picture in enc->cur (src/encoder.c:2631,
 * 2690), and notes that this encoder starts a picture of the given type.
 *
 * tcme_hook_motion_est receives every tc_motion_est call of encoder.c.  The
 * first call of a picture comes after estimate_global_mv
 * (src/encoder.c:2838-2841), so it declares the picture to the library with
 * the searches the encoder's call sites make at centres known at that point:
 *   P (src/encoder.c:892-1050): dpb[0] at the block position (the predictor
 *     site with a zero predictor, :926-928), dpb[1] likewise with
 *     multi-reference (:979-994), dpb[0] at the global-motion centre when
 *     |glob| > 16 quarter samples (:1000-1004), dpb[0] at the previous
 *     frame's displacement of the node's 8x8 cell (:1023-1035);
 *   B (src/encoder.c:795-852): every picture of the buffer at the block
 *     position (the forward and backward references are among them);
 *   v1 (src/encoder.c:1651-1701): nothing (its blocks are stack copies, and
 *     its calls go to the per-call drop-in).
 * Search ranges follow the preset (:801, :926) and so do the block sizes the
 * quadtree evaluates (:1328-1336).  Whatever the speculation gets wrong only
 * costs time: a call's answer is its own search, found or made.
 */
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include <pthread.h>

#include "tcodec_common.h"
#define TCME_HIDE_DROPIN
#include "tcme.h"

#define MAX_ENCODERS 16

typedef struct {
    tc_encoder_t *enc;
    tcme_frame *frame;
    int pending;                 /* a picture started and is not declared yet */
    tc_frame_type_t type;
    tcme_cell *grid;
    size_t grid_cap;
} hook_encoder;

static hook_encoder g_encs[MAX_ENCODERS];
static int g_nencs;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;

static void hook_die(const char *what)
{
    fprintf(stderr, "tcme hook: fatal: %s: %s\n", what, tcme_last_error());
    abort();
}

int tcme_hook_stats(tcme_frame_stats *out);

/* With TCME_HOOK_STATS set, one line of counters goes to stderr at exit. */
static void report(void)
{
    tcme_frame_stats s;
    tcme_hook_stats(&s);
    fprintf(stderr, "TCME_FRAME frames=%lld calls=%lld hits_spec=%lld hits_batch=%lld "
            "misses=%lld batch_calls=%lld uncached=%lld spec_calls=%lld begin_s=%.6f "
            "wait_s=%.6f miss_s=%.6f\n", (long long)s.frames, (long long)s.calls,
            (long long)s.hits_spec, (long long)s.hits_batch, (long long)s.misses,
            (long long)s.batch_calls, (long long)s.uncached, (long long)s.spec_calls,
            s.begin_s, s.wait_s, s.miss_s);
}

void tcme_hook_frame_start(tc_ratectl_t *rc, tc_frame_type_t type)
{
    tc_ratectl_frame_start(rc, type);
    tc_encoder_t *enc = (tc_encoder_t *)((char *)rc - offsetof(tc_encoder_t, rc));
    tcme_prepare();  /* the GPU runtime starts while the encoder works */
    pthread_mutex_lock(&g_mu);
    hook_encoder *h = NULL;
    for (int i = 0; i < g_nencs && !h; i++)
        if (g_encs[i].enc == enc) h = &g_encs[i];
    if (!h && g_nencs < MAX_ENCODERS) {
        if (g_nencs == 0 && getenv("TCME_HOOK_STATS")) atexit(report);
        h = &g_encs[g_nencs++];
        h->enc = enc;
    }
    if (h) {
        if (h->frame && tcme_frame_end(h->frame) != 0) hook_die("tcme_frame_end");
        h->pending = 1;
        h->type = type;
    }
    pthread_mutex_unlock(&g_mu);
}

/* Declares h's picture (called with g_mu held, at its first search). */
static void declare(hook_encoder *h)
{
    tc_encoder_t *enc = h->enc;
    if (!h->frame && tcme_frame_create(&h->frame) != 0) hook_die("tcme_frame_create");
    const uint8_t *refs[TC_REF_FRAMES];
    int strides[TC_REF_FRAMES];
    for (int i = 0; i < TC_REF_FRAMES; i++) {
        const tc_frame_buf_t *fr = enc->dpb[i].frame;
        refs[i] = fr ? fr->y : NULL;
        strides[i] = fr ? fr->stride_y : enc->cfg.width;
    }
    const int preset = enc->cfg.preset;
    const int fast = preset <= TC_PRESET_FAST;
    /* ultrafast evaluates 8x8 leaves only, fast 16x16 and 8x8 (:1328-1336) */
    const int sizes = preset == TC_PRESET_ULTRAFAST ? 0x1 : fast ? 0x3 : 0xF;
    tcme_spec specs[TC_REF_FRAMES + 3];
    int n = 0;
    if (enc->cfg.use_v2 && h->type == TC_FRAME_INTER && refs[0]) {
        const int sr = fast ? 16 : 24;
        const tcme_spec zero = {0, sr, sizes, TCME_CENTRE_OFFSET, 0, 0, NULL, 0, 0};
        specs[n++] = zero;
        const int multiref = preset >= TC_PRESET_MEDIUM &&
                             enc->cfg.profile >= TC_PROFILE_STREAMING_MAIN;
        if (multiref && refs[1]) {
            specs[n] = zero;
            specs[n++].ref = 1;
        }
        if (!fast && enc->glob_mv_valid &&
            (enc->glob_mv_x > 16 || enc->glob_mv_x < -16 ||
             enc->glob_mv_y > 16 || enc->glob_mv_y < -16)) {
            specs[n] = zero;
            specs[n].dx = enc->glob_mv_x;
            specs[n++].dy = enc->glob_mv_y;
        }
        if (!fast && enc->prev_mv_valid && enc->prev_mvgrid) {
            const size_t cells = (size_t)enc->prev_mvgrid_w * (size_t)enc->prev_mvgrid_h;
            if (cells > h->grid_cap) {
                free(h->grid);
                h->grid = (tcme_cell *)malloc(cells * sizeof(tcme_cell));
                if (!h->grid) hook_die("displacement grid");
                h->grid_cap = cells;
            }
            for (size_t i = 0; i < cells; i++) {
                const qt_mvcell_t *c = &enc->prev_mvgrid[i];
                h->grid[i].dx = c->dx;
                h->grid[i].dy = c->dy;
                h->grid[i].unavailable = c->intra ? 1 : 0;
                h->grid[i].reserved = 0;
            }
            specs[n] = zero;
            specs[n].mode = TCME_CENTRE_GRID;
            specs[n].grid = h->grid;
            specs[n].grid_w = enc->prev_mvgrid_w;
            specs[n++].grid_h = enc->prev_mvgrid_h;
        }
    } else if (enc->cfg.use_v2 && h->type == TC_FRAME_BIDIR) {
        const int sr = fast ? 16 : 32;
        for (int i = 0; i < TC_REF_FRAMES; i++)
            if (refs[i]) {
                const tcme_spec s = {i, sr, sizes, TCME_CENTRE_OFFSET, 0, 0, NULL, 0, 0};
                specs[n++] = s;
            }
    }
    if (n == 0) return;  /* nothing to search ahead: calls go to the per-call drop-in */
    if (tcme_frame_begin(h->frame, enc->cur->y, enc->cur->stride_y, enc->cfg.width,
                         enc->cfg.height, refs, strides, TC_REF_FRAMES, specs, n) != 0)
        hook_die("tcme_frame_begin");
}

tc_mv_s tcme_hook_motion_est(const tc_pixel_t *ref, int ref_stride, int ref_w, int ref_h,
                             const tc_pixel_t *orig, int orig_stride, int center_x,
                             int center_y, int blk_size, int search_range,
                             tc_sad_t *best_sad_out)
{
    tcme_frame *f = NULL;
    pthread_mutex_lock(&g_mu);
    for (int i = 0; i < g_nencs && !f; i++) {
        hook_encoder *h = &g_encs[i];
        int mine = 0;
        for (int k = 0; k < TC_REF_FRAMES; k++)
            mine |= h->enc->dpb[k].frame && h->enc->dpb[k].frame->y == ref;
        if (!mine) continue;
        if (h->pending) {
            declare(h);
            h->pending = 0;
        }
        f = h->frame;
    }
    pthread_mutex_unlock(&g_mu);
    if (!f)
        return tc_motion_est(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x,
                             center_y, blk_size, search_range, best_sad_out);
    int32_t sad = 0;
    const tcme_mv mv = tcme_frame_motion_est(f, ref, ref_stride, ref_w, ref_h, orig, orig_stride,
                                             center_x, center_y, blk_size, search_range, &sad);
    if (best_sad_out) *best_sad_out = sad;
    tc_mv_s out = {mv.x, mv.y};
    return out;
}

/* Counters of every encoder's frame object, summed (for reports). */
int tcme_hook_stats(tcme_frame_stats *out)
{
    tcme_frame_stats sum = {0};
    pthread_mutex_lock(&g_mu);
    for (int i = 0; i < g_nencs; i++) {
        tcme_frame_stats s;
        if (!g_encs[i].frame || tcme_frame_get_stats(g_encs[i].frame, &s) != 0) continue;
        sum.frames += s.frames;
        sum.spec_calls += s.spec_calls;
        sum.calls += s.calls;
        sum.hits_spec += s.hits_spec;
        sum.hits_batch += s.hits_batch;
        sum.misses += s.misses;
        sum.batch_calls += s.batch_calls;
        sum.uncached += s.uncached;
        sum.begin_s += s.begin_s;
        sum.wait_s += s.wait_s;
        sum.miss_s += s.miss_s;
    }
    pthread_mutex_unlock(&g_mu);
    *out = sum;
    return 0;
}
