/*
 * tcme_oracle.c - CPU restatement of TCodec's motion-estimation path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for libtcme.so, never shipped or
 * called by it.  Plain C, integer arithmetic only, written from the
 * reference's behaviour (file:line cited per function, paths relative to the
 * reference tree).
 *
 * Pinning, per function (tests/test_oracle_*.py):
 *   orc_sad, orc_sad_subpel, orc_motion_est, orc_inter_predict,
 *   orc_inter_predict_chroma   - against the compiled reference functions
 *                                (oracle/_ref/libtcref.so) on random and
 *                                edge inputs, and against golden call traces
 *                                recorded from the reference encoder
 *   orc_half_*, orc_qpel6      - closed forms (H.264 half-pel taps on constant
 *                                and linear ramps), brute-force bounds, and
 *                                through orc_inter_predict / orc_motion_est
 *   orc_mvp, orc_global_mv     - against the reference's static helpers,
 *                                exported by oracle/ref_expose.c
 */
#include "tcme_oracle.h"
#include <stddef.h>
#include <stdint.h>
#include <string.h>

static inline int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }
static inline int iabs(int v) { return v < 0 ? -v : v; }

/* Block SAD - src/motion.c:40-52. */
int32_t orc_sad(const uint8_t *a, int sa, const uint8_t *b, int sb, int n)
{
    int32_t s = 0;
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++) s += iabs((int)a[r * sa + c] - (int)b[r * sb + c]);
    return s;
}

/* Six-tap [1,-5,20,20,-5,1], +16, arithmetic >>5, no clip - src/motion.c:70-107. */
static inline int tap6(int p0, int p1, int p2, int p3, int p4, int p5)
{
    return (p0 + p5 - 5 * (p1 + p4) + 20 * (p2 + p3) + 16) >> 5;
}

/* Raw horizontal half-sample between (x,y) and (x+1,y): src/motion.c:70-81. */
int orc_half_h(const uint8_t *ref, int stride, int x, int y)
{
    const uint8_t *p = ref + (ptrdiff_t)y * stride + x;
    return tap6(p[-2], p[-1], p[0], p[1], p[2], p[3]);
}

/* Raw vertical half-sample between (x,y) and (x,y+1): src/motion.c:84-93. */
int orc_half_v(const uint8_t *ref, int stride, int x, int y)
{
    const uint8_t *p = ref + (ptrdiff_t)y * stride + x;
    return tap6(p[-2 * stride], p[-stride], p[0], p[stride], p[2 * stride], p[3 * stride]);
}

/* Raw centre half-sample: vertical taps over six raw horizontal half-samples,
 * src/motion.c:96-107. */
int orc_half_hv(const uint8_t *ref, int stride, int x, int y)
{
    int h[6];
    for (int k = 0; k < 6; k++) h[k] = orc_half_h(ref, stride, x, y - 2 + k);
    return tap6(h[0], h[1], h[2], h[3], h[4], h[5]);
}

/* Quarter-sample luma from nine anchors with raw half-samples and one final
 * clip, src/motion.c:193-260.  Row interpolation picks from the integer /
 * half / integer triple of each anchor row; the column step blends rows. */
int orc_qpel6(const uint8_t *ref, int stride, int x, int y, int dx, int dy)
{
    const uint8_t *p = ref + (ptrdiff_t)y * stride + x;
    int i00 = p[0], i10 = p[1], i01 = p[stride], i11 = p[stride + 1];
    int hb = orc_half_h(ref, stride, x, y);       /* half between i00,i10 */
    int he = orc_half_h(ref, stride, x, y + 1);   /* half between i01,i11 */
    int vc = orc_half_v(ref, stride, x, y);       /* half between i00,i01 */
    int vf = orc_half_v(ref, stride, x + 1, y);   /* half between i10,i11 */
    int cd = orc_half_hv(ref, stride, x, y);      /* centre */
    /* Each anchor row is a 3-point lattice L0, L1 (half), L2; phase dx picks
     * L0, avg(L0,L1), L1, avg(L1,L2). */
    int rows[3][3] = {{i00, hb, i10}, {vc, cd, vf}, {i01, he, i11}};
    int v[3];
    for (int r = 0; r < 3; r++) {
        const int *L = rows[r];
        switch (dx & 3) {
        case 0: v[r] = L[0]; break;
        case 1: v[r] = (L[0] + L[1] + 1) >> 1; break;
        case 2: v[r] = L[1]; break;
        default: v[r] = (L[1] + L[2] + 1) >> 1; break;
        }
    }
    int out;
    switch (dy & 3) {
    case 0: out = v[0]; break;
    case 1: out = (v[0] + v[1] + 1) >> 1; break;
    case 2: out = v[1]; break;
    default: out = (v[1] + v[2] + 1) >> 1; break;
    }
    return clip255(out);
}

/* Bilinear quarter-sample with a clip after each pass, src/motion.c:111-116
 * applied as in :291-297 and :330-336.  All operands are non-negative, so
 * C's "/4" equals ">>2". */
static inline int lerp4(int a, int b, int f) { return clip255((a * (4 - f) + b * f + 2) / 4); }

int orc_bilin(const uint8_t *ref, int stride, int x, int y, int dx, int dy)
{
    const uint8_t *p = ref + (ptrdiff_t)y * stride + x;
    int top = lerp4(p[0], p[1], dx);
    int bot = lerp4(p[stride], p[stride + 1], dx);
    return lerp4(top, bot, dy);
}

/* Bilinear sub-pel SAD, src/motion.c:308-341. */
int32_t orc_sad_subpel(const uint8_t *ref, int stride, int qx, int qy,
                       const uint8_t *orig, int ostride, int n)
{
    int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
    int32_t s = 0;
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            s += iabs((int)orig[r * ostride + c] - orc_bilin(ref, stride, fx + c, fy + r, dx, dy));
    return s;
}

/* Six-tap quarter-pel SAD used by the ME refinement, src/motion.c:463-471. */
int32_t orc_sad_qpel6(const uint8_t *ref, int stride, int qx, int qy,
                      const uint8_t *orig, int ostride, int n)
{
    int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
    int32_t s = 0;
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            s += iabs((int)orig[r * ostride + c] - orc_qpel6(ref, stride, fx + c, fy + r, dx, dy));
    return s;
}

/* ---- the search -------------------------------------------------------- */

typedef struct {
    const uint8_t *ref, *orig;
    int stride, ostride, w, h, n;
    orc_stats *st;
} me_ctx;

/* Integer candidate cost, src/motion.c:371-380: off-frame -> INT32_MAX.
 * Coordinates are taken as 64-bit: the reference forms them in int
 * (motion.c:401-402), which overflows for search ranges near INT_MAX; read
 * mathematically such a candidate is simply off the frame. */
static int32_t cost_int(const me_ctx *m, long long x, long long y)
{
    if (x < 0 || y < 0 || x + m->n > m->w || y + m->n > m->h) {
        if (m->st) m->st->int_oob++;
        return 0x7FFFFFFF;
    }
    if (m->st) m->st->int_evals++;
    return orc_sad(m->ref + (ptrdiff_t)y * m->stride + x, m->stride, m->orig, m->ostride, m->n);
}

/* Quarter-pel candidate cost, src/motion.c:455-475; returns 0 when skipped. */
static int cost_qpel(const me_ctx *m, int qx, int qy, int32_t *s)
{
    int fx = qx >> 2, fy = qy >> 2, n = m->n;
    if (fx < 0 || fy < 0 || fx + n + 1 > m->w || fy + n + 1 > m->h) {
        if (m->st) m->st->q_skip++;
        return 0;
    }
    if (fx >= 2 && fy >= 2 && fx + n + 2 < m->w && fy + n + 2 < m->h) {
        if (m->st) m->st->q6_evals++;
        *s = orc_sad_qpel6(m->ref, m->stride, qx, qy, m->orig, m->ostride, n);
    } else {
        if (m->st) m->st->qbil_evals++;
        *s = orc_sad_subpel(m->ref, m->stride, qx, qy, m->orig, m->ostride, n);
    }
    return 1;
}

/* Search patterns (offsets after the implicit {0,0}), src/motion.c:347-366. */
static const int8_t PAT_LARGE[6][2] = {{-2, -4}, {2, -4}, {-4, 0}, {4, 0}, {-2, 4}, {2, 4}};
static const int8_t PAT_SMALL[6][2] = {{-1, -2}, {1, -2}, {-2, 0}, {2, 0}, {-1, 2}, {1, 2}};
static const int8_t PAT_DIAM[4][2]  = {{0, -1}, {-1, 0}, {1, 0}, {0, 1}};
static const int8_t PAT_QPEL[8][2]  = {{1, 0}, {-1, 0}, {0, 1}, {0, -1},
                                       {1, 1}, {-1, 1}, {1, -1}, {-1, -1}};

/* One greedy pattern stage over full-pel positions: each pass visits the
 * pattern in table order around the *current* best (which moves on every
 * strict improvement); a pass with no improvement ends the stage.
 * src/motion.c:399-414, 416-431, 433-448. */
static void stage_int(const me_ctx *m, const int8_t (*pat)[2], int npat, int scale,
                      int passes, int *bx, int *by, int32_t *best)
{
    for (int it = 0; it < passes; it++) {
        int moved = 0;
        for (int k = 0; k < npat; k++) {
            long long x = *bx + (long long)pat[k][0] * scale;
            long long y = *by + (long long)pat[k][1] * scale;
            int32_t s = cost_int(m, x, y);
            if (s < *best) { *best = s; *bx = (int)x; *by = (int)y; moved = 1; }
        }
        if (!moved) return;
    }
}

/* Hierarchical hexagon search, src/motion.c:384-485.  Returns the absolute
 * quarter-pel position of the best block. */
orc_mv orc_motion_est(const uint8_t *ref, int stride, int w, int h,
                      const uint8_t *orig, int ostride, int cx, int cy,
                      int n, int sr, int32_t *sad_out, orc_stats *st)
{
    me_ctx m = {ref, orig, stride, ostride, w, h, n, st};
    /* Centre clamp as tc_clip(v, 0, W-n): low bound tested first. */
    cx = cx < 0 ? 0 : (cx > w - n ? w - n : cx);
    cy = cy < 0 ? 0 : (cy > h - n ? h - n : cy);
    int bx = cx, by = cy;
    int32_t best = cost_int(&m, bx, by);
    int scale = sr / 4 > 1 ? sr / 4 : 1;
    stage_int(&m, PAT_LARGE, 6, scale, 2, &bx, &by, &best);
    stage_int(&m, PAT_SMALL, 6, 1, 3, &bx, &by, &best);
    stage_int(&m, PAT_DIAM, 4, 1, 4, &bx, &by, &best);
    int qx = bx * 4, qy = by * 4;
    for (int k = 0; k < 8; k++) {
        int tx = qx + PAT_QPEL[k][0], ty = qy + PAT_QPEL[k][1];
        int32_t s;
        if (cost_qpel(&m, tx, ty, &s) && s < best) { best = s; qx = tx; qy = ty; }
    }
    if (sad_out) *sad_out = best;
    orc_mv r = {qx, qy};
    return r;
}

void orc_motion_est_batch(const uint8_t *ref, int rstride, int w, int h,
                          const uint8_t *cur, int cstride, const int32_t *calls,
                          int ncalls, int32_t *out, orc_stats *stats)
{
    for (int i = 0; i < ncalls; i++) {
        const int32_t *c = calls + 6 * i;
        int32_t sad;
        orc_mv mv = orc_motion_est(ref, rstride, w, h, cur + (ptrdiff_t)c[1] * cstride + c[0],
                                   cstride, c[2], c[3], c[4], c[5], &sad,
                                   stats ? stats + i : NULL);
        out[3 * i] = mv.x;
        out[3 * i + 1] = mv.y;
        out[3 * i + 2] = sad;
    }
}

/* Algorithmic bytes of one call: each in-frame integer SAD reads n^2 current
 * and n^2 reference samples, a 6-tap quarter-pel SAD n^2 + (n+5)^2, a bilinear
 * one n^2 + (n+1)^2. */
uint64_t orc_algo_bytes(const orc_stats *st, int n)
{
    uint64_t nn = (uint64_t)n * n;
    return (uint64_t)st->int_evals * 2 * nn +
           (uint64_t)st->q6_evals * (nn + (uint64_t)(n + 5) * (n + 5)) +
           (uint64_t)st->qbil_evals * (nn + (uint64_t)(n + 1) * (n + 1));
}

/* ---- motion compensation ----------------------------------------------- */

/* One luma prediction sample with the reference's edge ladder:
 * 6-tap when (x-2..x+3, y-2..y+3) is inside, else bilinear when (x..x+1,
 * y..y+1) is inside, else the nearest clamped sample, src/motion.c:269-304.
 * Inside the 6-tap region the h/v-only formulas of :133-191 agree with the
 * nine-anchor form after the final clip (they differ only where the
 * pre-clip sum is negative), so one formulation serves every phase. */
int orc_luma_px(const uint8_t *ref, int stride, int w, int h, int x, int y, int dx, int dy)
{
    if (x >= 2 && x + 3 < w && y >= 2 && y + 3 < h) {
        if (dx == 0 && dy == 0) return ref[(ptrdiff_t)y * stride + x];
        return orc_qpel6(ref, stride, x, y, dx, dy);
    }
    if (x >= 0 && y >= 0 && x + 1 < w && y + 1 < h) return orc_bilin(ref, stride, x, y, dx, dy);
    int nx = x + (dx >= 2 ? 1 : 0), ny = y + (dy >= 2 ? 1 : 0);
    nx = nx < 0 ? 0 : (nx > w - 1 ? w - 1 : nx);
    ny = ny < 0 ? 0 : (ny > h - 1 ? h - 1 : ny);
    return ref[(ptrdiff_t)ny * stride + nx];
}

/* Luma MC, src/motion.c:493-584: every branch there (row copy, whole-block
 * 6-tap, per-sample ladder) yields orc_luma_px per sample. */
void orc_inter_predict(const uint8_t *ref, int stride, int w, int h, int mvx, int mvy,
                       uint8_t *dst, int dstride, int n)
{
    int fx = mvx >> 2, fy = mvy >> 2, dx = mvx & 3, dy = mvy & 3;
    for (int r = 0; r < n; r++)
        for (int c = 0; c < n; c++)
            dst[r * dstride + c] = (uint8_t)orc_luma_px(ref, stride, w, h, fx + c, fy + r, dx, dy);
}

/* Chroma MC at 1/8 sample, edge-clamped bilinear, src/motion.c:597-632. */
void orc_inter_predict_chroma(const uint8_t *ref, int stride, int w, int h, int mvx,
                              int mvy, uint8_t *dst, int dstride, int n)
{
    int xi = mvx >> 3, yi = mvy >> 3, fx = mvx & 7, fy = mvy & 7;
    for (int r = 0; r < n; r++) {
        int y0 = yi + r, y1 = yi + r + 1;
        y0 = y0 < 0 ? 0 : (y0 > h - 1 ? h - 1 : y0);
        y1 = y1 < 0 ? 0 : (y1 > h - 1 ? h - 1 : y1);
        for (int c = 0; c < n; c++) {
            int x0 = xi + c, x1 = xi + c + 1;
            x0 = x0 < 0 ? 0 : (x0 > w - 1 ? w - 1 : x0);
            x1 = x1 < 0 ? 0 : (x1 > w - 1 ? w - 1 : x1);
            int a = ref[(ptrdiff_t)y0 * stride + x0], b = ref[(ptrdiff_t)y0 * stride + x1];
            int cc = ref[(ptrdiff_t)y1 * stride + x0], d = ref[(ptrdiff_t)y1 * stride + x1];
            int v = (8 - fx) * (8 - fy) * a + fx * (8 - fy) * b + (8 - fx) * fy * cc + fx * fy * d;
            dst[r * dstride + c] = (uint8_t)((v + 32) >> 6);
        }
    }
}

/* ---- search centres ------------------------------------------------------ */

/* Minimum-magnitude MV predictor over the 8x8-cell CTU grid (row stride 8),
 * src/encoder.c:405-427.  Neighbours: left, above, above-right (not in the
 * last column); intra cells are unavailable; missing ones are substituted. */
void orc_mvp(int cx, int cy, const int16_t *dx, const int16_t *dy, const uint8_t *intra,
             int *out_x, int *out_y)
{
    int ax = 0, ay = 0, bx = 0, by = 0, qx = 0, qy = 0;
    int has_a = 0, has_b = 0, has_c = 0;
    if (cx > 0 && !intra[cy * 8 + cx - 1]) { ax = dx[cy * 8 + cx - 1]; ay = dy[cy * 8 + cx - 1]; has_a = 1; }
    if (cy > 0 && !intra[(cy - 1) * 8 + cx]) { bx = dx[(cy - 1) * 8 + cx]; by = dy[(cy - 1) * 8 + cx]; has_b = 1; }
    if (cy > 0 && cx < 7 && !intra[(cy - 1) * 8 + cx + 1]) {
        qx = dx[(cy - 1) * 8 + cx + 1]; qy = dy[(cy - 1) * 8 + cx + 1]; has_c = 1;
    }
    if (!has_a && !has_b && !has_c) { *out_x = 0; *out_y = 0; return; }
    if (!has_a) { if (has_b) { ax = bx; ay = by; } else { ax = qx; ay = qy; } }
    if (!has_b) { if (has_a) { bx = ax; by = ay; } else { bx = qx; by = qy; } }
    if (!has_c) { if (has_a) { qx = ax; qy = ay; } else { qx = bx; qy = by; } }
    int64_t ma = (int64_t)ax * ax + (int64_t)ay * ay;
    int64_t mb = (int64_t)bx * bx + (int64_t)by * by;
    int64_t mc = (int64_t)qx * qx + (int64_t)qy * qy;
    if (mb < ma && mb <= mc) { *out_x = bx; *out_y = by; }
    else if (mc < ma && mc < mb) { *out_x = qx; *out_y = qy; }
    else { *out_x = ax; *out_y = ay; }
}

/* Global translation hint, src/encoder.c:171-202: two greedy passes over a
 * sparse sample lattice (every 8th pixel of every 8th row); pass 0 walks a
 * 9x9 grid of 8-px steps, pass 1 a 3x3 grid of 4-px steps; off-frame samples
 * cost 1020.  Output in quarter-pel. */
void orc_global_mv(const uint8_t *cur, int cstride, const uint8_t *ref, int rstride,
                   int w, int h, int *gx, int *gy)
{
    int cw = w / 4, ch = h / 4, ox = 0, oy = 0;
    int64_t best = (int64_t)1 << 60;
    for (int pass = 0; pass < 2; pass++) {
        int step = pass ? 1 : 2, rng = pass ? 1 : 8, moved = 0;
        for (int sy = -rng; sy <= rng; sy += step)
            for (int sx = -rng; sx <= rng; sx += step) {
                int64_t s = 0;
                for (int y = 0; y < ch; y += 2)
                    for (int x = 0; x < cw; x += 2) {
                        int rx = 4 * x + ox + 4 * sx, ry = 4 * y + oy + 4 * sy;
                        if (rx < 0 || ry < 0 || rx >= w || ry >= h) { s += 1020; continue; }
                        s += iabs((int)cur[(ptrdiff_t)(4 * y) * cstride + 4 * x] -
                                  (int)ref[(ptrdiff_t)ry * rstride + rx]);
                    }
                if (s < best) { best = s; ox += 4 * sx; oy += 4 * sy; moved = 1; }
            }
        if (!moved && pass) break;
    }
    *gx = ox * 4;
    *gy = oy * 4;
}

/* ---- ME-only predictor chain -------------------------------------------
 * The reference encoder centres each search at qt_mvp over a per-CTU 8x8
 * grid (src/encoder.c:405-427; the grid is reset to "unavailable" per CTU at
 * :1425-1428) that its RDO fills with each leaf's final displacement
 * (:1279-1288), visiting nodes in qt_split's decision order (:1292-1397: the
 * parent as a leaf, the four children as leaves, then, when the children win
 * (:1388, child <= parent), each child again as a subtree; nodes not wholly
 * in the frame are split without a search, :1299-1306).  The chain keeps that
 * order, that grid and that predictor and replaces the RDO with an ME-only
 * rule: a leaf searches every reference at ((mvp + 4 pos) >> 2) and keeps the
 * cheapest J = SAD + lambda * (se_bits(mvd.x) + se_bits(mvd.y)) (the first
 * reference on ties), mvd = mv - (mvp + 4 pos), se_bits as
 * src/bitstream.c:100-112; its cells take the chosen displacement mv - 4 pos
 * (as :1283-1285 stores mvp + mvd).  The predictor and the search are
 * function pointers so the tests can run the chain on the reference's own
 * qt_mvp and tc_motion_est. */

int orc_se_bits(int32_t v)
{
    uint32_t m = v > 0 ? (uint32_t)(2 * (int64_t)v - 1) : v < 0 ? (uint32_t)(-2 * (int64_t)v) : 0u;
    uint64_t code = (uint64_t)m + 1;
    int bits = 0;
    while (code) { bits++; code >>= 1; }
    return 2 * bits - 1;
}

static orc_mv chain_me_default(const uint8_t *ref, int stride, int w, int h, const uint8_t *orig,
                               int ostride, int cx, int cy, int n, int sr, int32_t *sad)
{
    return orc_motion_est(ref, stride, w, h, orig, ostride, cx, cy, n, sr, sad, NULL);
}

typedef struct {
    const uint8_t *const *refs;
    int nrefs, rstride, w, h, sr, lambda;
    const uint8_t *cur;
    int cstride;
    orc_mvp_fn mvp;
    orc_me_fn me;
    int16_t dx[64], dy[64];
    uint8_t intra[64];
    int ctu, ctu_x, ctu_y;
    int32_t *out;
    int n, cap;
} chain_ctx;

static int64_t chain_leaf(chain_ctx *c, int depth, int cx, int cy)
{
    const int cu = 64 >> depth, px = c->ctu_x + 8 * cx, py = c->ctu_y + 8 * cy;
    int mx, my;
    c->mvp(cx, cy, c->dx, c->dy, c->intra, &mx, &my);
    const int ccx = (mx + 4 * px) >> 2, ccy = (my + 4 * py) >> 2;
    int64_t best = INT64_MAX;
    int bdx = 0, bdy = 0;
    for (int r = 0; r < c->nrefs; r++) {
        int32_t sad = 0;
        orc_mv mv = c->me(c->refs[r], c->rstride, c->w, c->h,
                          c->cur + (ptrdiff_t)py * c->cstride + px, c->cstride, ccx, ccy, cu,
                          c->sr, &sad);
        const int64_t j = (int64_t)sad + (int64_t)c->lambda *
                          (orc_se_bits(mv.x - (mx + 4 * px)) + orc_se_bits(mv.y - (my + 4 * py)));
        if (c->n < c->cap) {
            int32_t *row = c->out + (ptrdiff_t)c->n * ORC_CHAIN_ROW;
            row[0] = c->ctu; row[1] = r; row[2] = px; row[3] = py; row[4] = ccx; row[5] = ccy;
            row[6] = cu; row[7] = mv.x; row[8] = mv.y; row[9] = sad;
        }
        c->n++;
        if (j < best) { best = j; bdx = mv.x - 4 * px; bdy = mv.y - 4 * py; }
    }
    const int cs = cu / 8;
    for (int gy = 0; gy < cs; gy++)
        for (int gx = 0; gx < cs; gx++) {
            const int i = (cy + gy) * 8 + cx + gx;
            c->dx[i] = (int16_t)bdx;
            c->dy[i] = (int16_t)bdy;
            c->intra[i] = 0;
        }
    return best;
}

static void chain_save(const chain_ctx *c, int cx, int cy, int cs, int16_t *dx, int16_t *dy,
                       uint8_t *in)
{
    for (int gy = 0; gy < cs; gy++)
        for (int gx = 0; gx < cs; gx++) {
            const int i = (cy + gy) * 8 + cx + gx, k = gy * cs + gx;
            dx[k] = c->dx[i]; dy[k] = c->dy[i]; in[k] = c->intra[i];
        }
}

static void chain_load(chain_ctx *c, int cx, int cy, int cs, const int16_t *dx, const int16_t *dy,
                       const uint8_t *in)
{
    for (int gy = 0; gy < cs; gy++)
        for (int gx = 0; gx < cs; gx++) {
            const int i = (cy + gy) * 8 + cx + gx, k = gy * cs + gx;
            c->dx[i] = dx[k]; c->dy[i] = dy[k]; c->intra[i] = in[k];
        }
}

static void chain_split(chain_ctx *c, int depth, int cx, int cy)
{
    const int cu = 64 >> depth, px = c->ctu_x + 8 * cx, py = c->ctu_y + 8 * cy;
    const int in_frame = px + cu <= c->w && py + cu <= c->h;
    const int half = cu / 16;  /* child offset in cells */
    if (!in_frame && cu > 8) {
        for (int q = 0; q < 4; q++) chain_split(c, depth + 1, cx + (q & 1) * half, cy + (q >> 1) * half);
        return;
    }
    if (!in_frame) return;
    if (cu == 8) {
        chain_leaf(c, depth, cx, cy);
        return;
    }
    const int cs = cu / 8;
    int16_t pdx[64], pdy[64], qdx[64], qdy[64];
    uint8_t pin[64], qin[64];
    chain_save(c, cx, cy, cs, pdx, pdy, pin);
    const int64_t parent = chain_leaf(c, depth, cx, cy);
    chain_save(c, cx, cy, cs, qdx, qdy, qin);
    chain_load(c, cx, cy, cs, pdx, pdy, pin);
    int64_t child = 0;
    for (int q = 0; q < 4; q++) child += chain_leaf(c, depth + 1, cx + (q & 1) * half, cy + (q >> 1) * half);
    if (child <= parent) {
        chain_load(c, cx, cy, cs, pdx, pdy, pin);
        for (int q = 0; q < 4; q++) chain_split(c, depth + 1, cx + (q & 1) * half, cy + (q >> 1) * half);
    } else {
        chain_load(c, cx, cy, cs, qdx, qdy, qin);
    }
}

int orc_chain_frame(const uint8_t *const *refs, int nrefs, int rstride, const uint8_t *cur,
                    int cstride, int w, int h, int sr, int lambda, orc_mvp_fn mvp, orc_me_fn me,
                    int32_t *out, int cap)
{
    chain_ctx c;
    memset(&c, 0, sizeof c);
    c.refs = refs; c.nrefs = nrefs; c.rstride = rstride; c.w = w; c.h = h; c.sr = sr;
    c.lambda = lambda; c.cur = cur; c.cstride = cstride;
    c.mvp = mvp ? mvp : orc_mvp;
    c.me = me ? me : chain_me_default;
    c.out = out; c.cap = cap;
    const int cols = (w + 63) / 64, rows = (h + 63) / 64;
    for (int ry = 0; ry < rows; ry++)
        for (int rx = 0; rx < cols; rx++) {
            c.ctu = ry * cols + rx;
            c.ctu_x = rx * 64;
            c.ctu_y = ry * 64;
            for (int i = 0; i < 64; i++) { c.dx[i] = 0; c.dy[i] = 0; c.intra[i] = 1; }
            chain_split(&c, 0, 0, 0);
        }
    return c.n;
}
