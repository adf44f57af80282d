/*
 * ref_expose.c - checker-only translation unit.
 *
 * Compiles the reference encoder (src/encoder.c, included from where it lies
 * under the reference tree; nothing is copied) and adds two exported wrappers
 * around its file-static ME helpers, so the tests can pin the oracle's
 * restatement of them against the real code:
 *
 *   refx_qt_mvp        -> qt_mvp              (src/encoder.c:405-427)
 *   refx_qt_mvp_v      -> the same, void (a function pointer for the oracle)
 *   refx_global_mv     -> estimate_global_mv  (src/encoder.c:171-202)
 *
 * and one loop over the reference's public tc_motion_est (src/motion.c:384),
 * so that a CPU baseline can hand a thread many calls in one foreign call:
 *
 *   refx_motion_est_batch
 *
 * Built only into oracle/_ref/libtcref.so by oracle/Makefile.
 */
#include "encoder.c"

int refx_qt_mvp(int cx, int cy, const int16_t *dx, const int16_t *dy,
                const uint8_t *intra, int *out_x, int *out_y)
{
    qt_mvcell_t grid[TC_MVGRID_STRIDE * TC_MVGRID_STRIDE];
    for (int i = 0; i < TC_MVGRID_STRIDE * TC_MVGRID_STRIDE; i++) {
        grid[i].dx = dx[i];
        grid[i].dy = dy[i];
        grid[i].intra = intra[i];
    }
    tc_mv_s m = qt_mvp(NULL, cx, cy, grid);
    *out_x = m.x;
    *out_y = m.y;
    return 0;
}

/* The same with orc_mvp_fn's exact type (void), so the oracle's chain can
 * call the reference's predictor through a function pointer. */
void refx_qt_mvp_v(int cx, int cy, const int16_t *dx, const int16_t *dy, const uint8_t *intra,
                   int *out_x, int *out_y)
{
    refx_qt_mvp(cx, cy, dx, dy, intra, out_x, out_y);
}

int refx_global_mv(const uint8_t *cur, int cur_stride,
                   const uint8_t *ref, int ref_stride, int w, int h,
                   int *gx, int *gy)
{
    tc_encoder_t *enc = (tc_encoder_t *)calloc(1, sizeof(*enc));
    if (!enc) return -1;
    tc_frame_buf_t curf, reff;
    memset(&curf, 0, sizeof(curf));
    memset(&reff, 0, sizeof(reff));
    curf.y = (tc_pixel_t *)cur;  curf.stride_y = cur_stride; curf.width = w; curf.height = h;
    reff.y = (tc_pixel_t *)ref;  reff.stride_y = ref_stride; reff.width = w; reff.height = h;
    enc->cfg.width = w;
    enc->cfg.height = h;
    enc->cur = &curf;
    estimate_global_mv(enc, &reff);
    *gx = enc->glob_mv_x;
    *gy = enc->glob_mv_y;
    free(enc);
    return 0;
}

/* calls: n rows of (px, py, cx, cy, blk, sr); out: n rows of (mv_x, mv_y,
 * sad), each exactly tc_motion_est's result for the block at (px, py) of cur. */
int refx_motion_est_batch(const uint8_t *ref, int ref_stride, int w, int h,
                          const uint8_t *cur, int cur_stride, const int32_t *calls, int n,
                          int32_t *out)
{
    for (int i = 0; i < n; i++) {
        const int32_t *c = calls + 6 * i;
        tc_sad_t sad = 0;
        tc_mv_s mv = tc_motion_est(ref, ref_stride, w, h,
                                   cur + (size_t)c[1] * cur_stride + c[0], cur_stride,
                                   c[2], c[3], c[4], c[5], &sad);
        out[3 * i] = mv.x;
        out[3 * i + 1] = mv.y;
        out[3 * i + 2] = sad;
    }
    return 0;
}
