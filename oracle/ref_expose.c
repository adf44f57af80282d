/*
 * ref_expose.c - checker-only translation unit.
 *
 * Compiles the reference encoder (src/encoder.c, included from where it lies
 * under the reference tree; nothing is copied) and adds two exported wrappers
 * around its file-static ME helpers, so the tests can pin the oracle's
 * restatement of them against the real code:
 *
 *   refx_qt_mvp        -> qt_mvp              (src/encoder.c:405-427)
 *   refx_global_mv     -> estimate_global_mv  (src/encoder.c:171-202)
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
