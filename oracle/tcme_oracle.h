/*
 * tcme_oracle.h - CPU restatement of TCodec's motion-estimation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker the GPU backend is compared
 * against; only the tests, the smoke check and the benchmark's CPU-baseline
 * leg may load it.  The product (libtcme.so) never links or calls it.
 *
 * Every function names the reference lines it restates.  Pinning (what makes
 * the restatement trustworthy) is done by the tests: against the reference
 * compiled from its own sources (oracle/_ref/libtcref.so), against committed
 * encoder call traces in tests/golden/, and against closed-form and
 * brute-force properties.
 */
#ifndef TCME_ORACLE_H
#define TCME_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { int32_t x, y; } orc_mv;

/* Candidate evaluations in the order the sequential reference performs them
 * (src/motion.c:384-485).  Out-of-frame integer candidates and skipped
 * quarter-pel candidates read no pixels. */
typedef struct {
    int32_t int_evals;    /* in-frame integer SADs, centre included */
    int32_t int_oob;      /* integer candidates rejected by the frame test */
    int32_t q6_evals;     /* quarter-pel candidates priced with the 6-tap path */
    int32_t qbil_evals;   /* quarter-pel candidates priced with the bilinear path */
    int32_t q_skip;       /* quarter-pel candidates skipped at the frame edge */
} orc_stats;

int32_t orc_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int n);
int     orc_half_h(const uint8_t *ref, int stride, int x, int y);
int     orc_half_v(const uint8_t *ref, int stride, int x, int y);
int     orc_half_hv(const uint8_t *ref, int stride, int x, int y);
int     orc_qpel6(const uint8_t *ref, int stride, int x, int y, int dx, int dy);
int     orc_bilin(const uint8_t *ref, int stride, int x, int y, int dx, int dy);
int32_t orc_sad_subpel(const uint8_t *ref, int stride, int qx, int qy,
                       const uint8_t *orig, int ostride, int n);
int32_t orc_sad_qpel6(const uint8_t *ref, int stride, int qx, int qy,
                      const uint8_t *orig, int ostride, int n);
orc_mv  orc_motion_est(const uint8_t *ref, int stride, int w, int h,
                       const uint8_t *orig, int ostride, int cx, int cy,
                       int n, int sr, int32_t *sad_out, orc_stats *st);
/* Many calls over one reference/current pair: calls[i] = {px, py, cx, cy, n, sr};
 * out[i] = {mv_x, mv_y, sad}; stats may be NULL. */
void    orc_motion_est_batch(const uint8_t *ref, int rstride, int w, int h,
                             const uint8_t *cur, int cstride, const int32_t *calls,
                             int ncalls, int32_t *out, orc_stats *stats);
uint64_t orc_algo_bytes(const orc_stats *st, int n);
int     orc_luma_px(const uint8_t *ref, int stride, int w, int h, int x, int y,
                    int dx, int dy);
void    orc_inter_predict(const uint8_t *ref, int stride, int w, int h, int mvx,
                          int mvy, uint8_t *dst, int dstride, int n);
void    orc_inter_predict_chroma(const uint8_t *ref, int stride, int w, int h,
                                 int mvx, int mvy, uint8_t *dst, int dstride, int n);
void    orc_mvp(int cx, int cy, const int16_t *dx, const int16_t *dy,
                const uint8_t *intra, int *out_x, int *out_y);
void    orc_global_mv(const uint8_t *cur, int cstride, const uint8_t *ref, int rstride,
                      int w, int h, int *gx, int *gy);

/* ME-only predictor chain over one frame (restated rules in tcme_oracle.c):
 * rows of ORC_CHAIN_ROW int32 {ctu, ref, px, py, cx, cy, n, mv_x, mv_y, sad},
 * CTUs in raster order, each CTU's calls in the order they are made.  mvp and
 * me may be NULL (orc_mvp, orc_motion_est) or the reference's own functions.
 * Returns the call count (rows past cap are counted, not written). */
#define ORC_CHAIN_ROW 10
typedef void (*orc_mvp_fn)(int cx, int cy, const int16_t *dx, const int16_t *dy,
                           const uint8_t *intra, int *out_x, int *out_y);
typedef orc_mv (*orc_me_fn)(const uint8_t *ref, int stride, int w, int h, const uint8_t *orig,
                            int ostride, int cx, int cy, int n, int sr, int32_t *sad);
int     orc_se_bits(int32_t v);
int     orc_chain_frame(const uint8_t *const *refs, int nrefs, int rstride, const uint8_t *cur,
                        int cstride, int w, int h, int sr, int lambda, orc_mvp_fn mvp,
                        orc_me_fn me, int32_t *out, int cap);

#ifdef __cplusplus
}
#endif
#endif
