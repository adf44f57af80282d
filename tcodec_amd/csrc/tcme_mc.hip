// gfx950 kernels for TCodec's motion compensation (the consumers of each
// search result):
//
//   mc_luma_kernel    tc_inter_predict, reference src/motion.c:493-584, with
//                     get_interp_luma's per-sample ladder (src/motion.c:269-304)
//   mc_chroma_kernel  tc_inter_predict_chroma, src/motion.c:597-632
//
// Each XCD works through a contiguous share of the list, a chunk per
// workgroup: small interior blocks are cut into column tasks that the
// workgroup's lanes share evenly, large luma blocks into row bands, one per
// wavefront.  Every branch of the reference's luma predictor (row copy
// for an in-frame integer MV, whole-block six-tap, per-sample six-tap /
// bilinear / nearest ladder) gives, sample for sample, the ladder's value, so
// the kernel computes the
// ladder directly: six-tap where the sample's 6x6 apron is inside the frame,
// bilinear where its 2x2 neighbourhood is, nearest (clamped) otherwise.  The
// six-tap value comes from the anchor plane (sample, hh, vv, cc per position)
// read as the half-sample lattice M[Y][X]: with X0 = q >> 1, X1 = (q+1) >> 1
// per axis at quarter-pel q,
//   v = clip((avg(M[Y0][X0], M[Y0][X1]) + avg(M[Y1][X0], M[Y1][X1]) + 1) >> 1),
// avg(a,b) = (a+b+1) >> 1 on raw values (src/motion.c:193-260; the h- and
// v-only paths of src/motion.c:133-191 agree after their final clip).
#include <algorithm>

#include "tcme_internal.h"

namespace tcme {
namespace {

typedef const __attribute__((address_space(1))) uint8_t *gptr8;
typedef const __attribute__((address_space(1))) uint64_t *gptr64;
typedef __attribute__((address_space(1))) uint8_t *gout8;

__device__ __forceinline__ int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

// Bilinear quarter-pel step, rounded per pass (src/motion.c:111-116).
__device__ __forceinline__ int lerp4(int a, int b, int f) { return (a * (4 - f) + b * f + 2) >> 2; }

// 16-bit field f (0 sample, 1 hh, 2 vv, 3 cc) of an anchor, sign-extended.
__device__ __forceinline__ int field(uint64_t a, int f)
{
    return static_cast<int>(static_cast<int16_t>(static_cast<uint16_t>(a >> (16 * f))));
}

// The six-tap value at fraction (dx,dy) from the sample's anchor a00 and its
// right (a10), lower (a01) and lower-right (a11) neighbours; a10 is only used
// at dx = 3, a01 at dy = 3, a11 at both, and may be anything otherwise.  With
// X0 = (4x+dx) >> 1 and X1 = (4x+dx+1) >> 1 the lattice columns lie in anchor
// column x, except X1 = 2x+2 (anchor x+1) at dx = 3; rows likewise.
__device__ __forceinline__ int lattice_px(uint64_t a00, uint64_t a10, uint64_t a01, uint64_t a11,
                                          int dx, int dy)
{
    // lattice (Y, X) for Y in {Y0, Y1}, X in {X0, X1}: field (Y&1)*2 + (X&1)
    const int X0 = dx >> 1, X1 = (dx + 1) >> 1, Y0 = dy >> 1, Y1 = (dy + 1) >> 1;
    const int f00 = (Y0 & 1) * 2 + (X0 & 1), f01 = (Y0 & 1) * 2 + (X1 & 1);
    const int f10 = (Y1 & 1) * 2 + (X0 & 1), f11 = (Y1 & 1) * 2 + (X1 & 1);
    const uint64_t r0c1 = X1 == 2 ? a10 : a00;   // row Y0: anchor of column X1
    const uint64_t r1c0 = Y1 == 2 ? a01 : a00;   // row Y1: anchor of column X0
    const uint64_t r1c1 = Y1 == 2 ? (X1 == 2 ? a11 : a01) : (X1 == 2 ? a10 : a00);
    const int top = (field(a00, f00) + field(r0c1, f01) + 1) >> 1;
    const int bot = (field(r1c0, f10) + field(r1c1, f11) + 1) >> 1;
    return clip255((top + bot + 1) >> 1);
}

// get_interp_luma (src/motion.c:269-304) at integer (x,y), fraction (dx,dy):
// a sample reads its own anchor, and its right / lower neighbours only at
// phase 3.
__device__ __forceinline__ int luma_px(const PicDev &R, int x, int y, int dx, int dy)
{
    const gptr8 Y = (gptr8)R.y;
    if (x >= 2 && x + 3 < R.w && y >= 2 && y + 3 < R.h) {
        if (dx == 0 && dy == 0) return Y[static_cast<size_t>(y) * R.pitch + x];
        const gptr64 an = (gptr64)R.an + static_cast<size_t>(y) * R.apitch + x;
        const uint64_t a00 = an[0];
        const uint64_t a10 = dx == 3 ? an[1] : a00;
        const uint64_t a01 = dy == 3 ? an[R.apitch] : a00;
        const uint64_t a11 = dx == 3 && dy == 3 ? an[R.apitch + 1] : a00;
        return lattice_px(a00, a10, a01, a11, dx, dy);
    }
    if (x >= 0 && y >= 0 && x + 1 < R.w && y + 1 < R.h) {
        const gptr8 p = Y + static_cast<size_t>(y) * R.pitch + x;
        const int top = lerp4(p[0], p[1], dx);
        const int bot = lerp4(p[R.pitch], p[R.pitch + 1], dx);
        return lerp4(top, bot, dy);
    }
    int nx = x + (dx >= 2 ? 1 : 0), ny = y + (dy >= 2 ? 1 : 0);
    nx = nx < 0 ? 0 : (nx > R.w - 1 ? R.w - 1 : nx);
    ny = ny < 0 ? 0 : (ny > R.h - 1 ? R.h - 1 : ny);
    return Y[static_cast<size_t>(ny) * R.pitch + nx];
}

// Validates a descriptor against the picture table and the output buffer.
// R holds the table entry of picture `cached` (-1: none) and is reloaded only
// when the descriptor names another picture (the table does not change
// during a launch).
__device__ __forceinline__ int check_pred(const PicDev *table, int npics, const tcme_pred &p,
                                          int64_t out_bytes, bool need_anchors, PicDev &R,
                                          int &cached)
{
    if (p.pic < 0 || p.pic >= npics) return TCME_E_PIC;
    if (p.pic != cached) {
        R = table[p.pic];
        cached = p.pic;
    }
    if (!R.y || (need_anchors && !R.an)) return TCME_E_PIC;
    if (p.blk < 1 || p.blk > kMaxBlk) return TCME_E_BLK;
    if (!span_fits(p.out_off, p.blk, p.blk, p.out_stride, out_bytes)) return TCME_E_ARG;
    return 0;
}

// Lane layout of one wavefront over an n x n block (1 <= n <= 64): lanes
// take column lane % n of rows lane / n, lane / n + 64 / n, ...; lanes past
// the last whole row of a pass idle.
struct BlockLanes {
    int c, r0, dr;
    __device__ BlockLanes(int n, int lane) : c(lane % n), r0(lane / n), dr(64 / n)
    {
        if (r0 >= dr) r0 = n;  // idle lane
    }
};

// lattice_px at a fraction known at compile time: each field is a fixed
// half of one dword, and only the neighbours phase 3 needs are read.
template <int DX, int DY>
__device__ __forceinline__ int lattice_fixed(uint64_t a00, uint64_t a10, uint64_t a01,
                                             uint64_t a11)
{
    constexpr int X0 = DX >> 1, X1 = (DX + 1) >> 1, Y0 = DY >> 1, Y1 = (DY + 1) >> 1;
    constexpr int f00 = (Y0 & 1) * 2 + (X0 & 1), f01 = (Y0 & 1) * 2 + (X1 & 1);
    constexpr int f10 = (Y1 & 1) * 2 + (X0 & 1), f11 = (Y1 & 1) * 2 + (X1 & 1);
    const uint64_t r0c1 = X1 == 2 ? a10 : a00;
    const uint64_t r1c0 = Y1 == 2 ? a01 : a00;
    const uint64_t r1c1 = Y1 == 2 ? (X1 == 2 ? a11 : a01) : (X1 == 2 ? a10 : a00);
    const int top = (field(a00, f00) + field(r0c1, f01) + 1) >> 1;
    const int bot = (field(r1c0, f10) + field(r1c1, f11) + 1) >> 1;
    return clip255((top + bot + 1) >> 1);
}

// Rows r0, r0 + dr, ... < n of column (lane) of an interior block at fraction
// (DX,DY), K rows per lane at a time, loads first.  a0 points at the anchor of
// the lane's sample in the block's first row, o at its output.  Rows past the
// block re-read its last row and are not stored.
template <int DX, int DY, int K>
__device__ __forceinline__ void mc_interior(gptr64 a0, int ap, gout8 o, int ostride, int r0,
                                            int dr, int n)
{
    for (int rb = r0; rb < n; rb += K * dr) {
        uint64_t a00[K], a10[K], a01[K], a11[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int r = min(rb + k * dr, n - 1);
            const gptr64 e = a0 + static_cast<size_t>(r) * ap;
            a00[k] = e[0];
            a10[k] = DX == 3 ? e[1] : 0;
            a01[k] = DY == 3 ? e[ap] : 0;
            a11[k] = DX == 3 && DY == 3 ? e[ap + 1] : 0;
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int r = rb + k * dr;
            if (r < n)
                o[static_cast<int64_t>(r) * ostride] =
                    static_cast<uint8_t>(lattice_fixed<DX, DY>(a00[k], a10[k], a01[k], a11[k]));
        }
    }
}

// The same for an integer MV: the samples themselves.
template <int K>
__device__ __forceinline__ void mc_copy(gptr8 y0, int pitch, gout8 o, int ostride, int r0, int dr,
                                        int n)
{
    for (int rb = r0; rb < n; rb += K * dr) {
        uint8_t v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = y0[static_cast<size_t>(min(rb + k * dr, n - 1)) * pitch];
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (rb + k * dr < n) o[static_cast<int64_t>(rb + k * dr) * ostride] = v[k];
    }
}

template <int K>
__device__ __forceinline__ void mc_interior_k(int dx, int dy, gptr64 a0, int ap, gout8 o,
                                              int ostride, int r0, int dr, int n)
{
    switch ((dy << 2) | dx) {
#define TCME_MC_CASE(DX, DY) \
    case ((DY) << 2) | (DX): mc_interior<DX, DY, K>(a0, ap, o, ostride, r0, dr, n); break;
    TCME_MC_CASE(1, 0) TCME_MC_CASE(2, 0) TCME_MC_CASE(3, 0)
    TCME_MC_CASE(0, 1) TCME_MC_CASE(1, 1) TCME_MC_CASE(2, 1) TCME_MC_CASE(3, 1)
    TCME_MC_CASE(0, 2) TCME_MC_CASE(1, 2) TCME_MC_CASE(2, 2) TCME_MC_CASE(3, 2)
    TCME_MC_CASE(0, 3) TCME_MC_CASE(1, 3) TCME_MC_CASE(2, 3) TCME_MC_CASE(3, 3)
#undef TCME_MC_CASE
    default: break;
    }
}

// Rows [rs, re) of one predicted block by the whole wavefront (the block
// index is wave-uniform, so the descriptor and the picture-table entry are
// scalar loads, and the entry is reused while the picture stays the same).
// Per-sample selects of the fraction made the sample loop issue-bound, so an
// interior block - every sample's 6x6 apron inside the frame, where the
// ladder is the six-tap value, or the sample itself at an integer MV - runs
// code specialised for its fraction, up to 4 rows per lane at a time, loads
// first (8 rows took 84 VGPRs, 5 wavefronts per SIMD, and was slower).
// Other blocks take the ladder sample by sample.  The status is written by
// the wavefront that takes row 0.
__device__ __forceinline__ void luma_rows(const PicDev *__restrict__ table, int npics,
                                          const tcme_pred &p, int64_t i, int rs, int re,
                                          uint8_t *__restrict__ out, int64_t out_bytes,
                                          int32_t *__restrict__ status, PicDev &R, int &cached,
                                          int lane)
{
    const int st = check_pred(table, npics, p, out_bytes, true, R, cached);
    if (status && lane == 0 && rs == 0) status[i] = st;
    if (st) return;
    re = re < p.blk ? re : p.blk;
    if (rs >= re) return;
    const int fx = p.mv_x >> 2, fy = p.mv_y >> 2, dx = p.mv_x & 3, dy = p.mv_y & 3;
    const BlockLanes bl(p.blk, lane);
    const bool inner = fx >= 2 && fy >= 2 && fx + p.blk + 3 <= R.w && fy + p.blk + 3 <= R.h;
    if (!inner) {
        const gout8 o = (gout8)out + p.out_off;
        for (int r = rs + bl.r0; r < re; r += bl.dr)
            o[static_cast<int64_t>(r) * p.out_stride + bl.c] =
                static_cast<uint8_t>(luma_px(R, fx + bl.c, fy + r, dx, dy));
        return;
    }
    const int nr = re - rs;                          // rows of the band
    const int rows = (nr + bl.dr - 1) / bl.dr;        // per lane; wave-uniform
    const gout8 oc = (gout8)out + p.out_off + static_cast<int64_t>(rs) * p.out_stride + bl.c;
    if ((dx | dy) == 0) {
        const gptr8 y0 = (gptr8)R.y + static_cast<size_t>(fy + rs) * R.pitch + fx + bl.c;
        if (rows >= 4) mc_copy<4>(y0, R.pitch, oc, p.out_stride, bl.r0, bl.dr, nr);
        else mc_copy<1>(y0, R.pitch, oc, p.out_stride, bl.r0, bl.dr, nr);
        return;
    }
    const gptr64 a0 = (gptr64)R.an + static_cast<size_t>(fy + rs) * R.apitch + fx + bl.c;
    if (rows >= 4) mc_interior_k<4>(dx, dy, a0, R.apitch, oc, p.out_stride, bl.r0, bl.dr, nr);
    else mc_interior_k<1>(dx, dy, a0, R.apitch, oc, p.out_stride, bl.r0, bl.dr, nr);
}

typedef unsigned long long u64x2 __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(1))) u64x2 *gptr64x2;

__device__ __forceinline__ int shfl32(int v, int src)
{
    return __builtin_amdgcn_ds_bpermute(src << 2, v);
}
__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src)
{
    const uint32_t lo = static_cast<uint32_t>(shfl32(static_cast<int>(v), src));
    const uint32_t hi = static_cast<uint32_t>(shfl32(static_cast<int>(v >> 32), src));
    return static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

// Rows [r0, r0 + 8) of column c of an interior block (every sample's six-tap
// apron inside the frame) at quarter-pel (mvx, mvy), anchors an / apitch,
// output o (the block's first row) with row stride os: the anchor pairs
// (x, x+1) of the 9 rows are loaded first, then each sample is the lattice
// formula with the anchor fields selected at run time (the fraction is held
// per lane).
__device__ __forceinline__ void column8(const uint64_t *an, int apitch, int mvx, int mvy, int c,
                                        int r0, uint8_t *o, int os)
{
    const int fx = mvx >> 2, fy = mvy >> 2, dx = mvx & 3, dy = mvy & 3;
    const gptr64 a = (gptr64)an + static_cast<size_t>(fy + r0) * apitch + fx + c;
    const size_t ap = static_cast<size_t>(apitch);
    u64x2 e[9];
#pragma unroll
    for (int r = 0; r < 9; ++r) e[r] = *(gptr64x2)(a + r * ap);
    const gout8 oc = (gout8)o + static_cast<int64_t>(r0) * os + c;
#pragma unroll
    for (int r = 0; r < 8; ++r)
        oc[static_cast<int64_t>(r) * os] =
            static_cast<uint8_t>(lattice_px(e[r].x, e[r].y, e[r + 1].x, e[r + 1].y, dx, dy));
}

// Workgroup b (4 wavefronts) takes a contiguous chunk of the list; the
// chunks of each XCD are contiguous (workgroup b runs on XCD b % 8), so the
// blocks of one CTU, whose quadtree levels read the same anchors, meet in
// one L2.  The chunk is taken 64 descriptors at a time, lane k of every
// wavefront holding descriptor k:
//  * interior 8x8 and 16x16 blocks (every sample's six-tap apron inside the
//    frame) are cut into tasks of one column x 8 rows (8 per 8x8 block, 32
//    per 16x16); the tasks of the 64 descriptors are numbered in a row
//    (prefix sum over the lanes) and the workgroup's 256 lanes take 256 at a
//    time, each finding its block by a binary search over the prefix sums
//    and the block's fields by lane permutes, then loading its 9 anchor rows
//    at once (column8): a pass is one round trip to memory and every
//    wavefront gets the same share (one wavefront per block spent 2-4
//    dependent rounds of loads on each);
//  * 32x32 and 64x64 blocks are shared by the four wavefronts, a quarter of
//    the rows each, with code specialised for the block's fraction (one
//    wavefront alone on a 64x64 block made 16 dependent rounds of loads; as
//    column tasks, with the fraction selected per sample, they cost 2.5x
//    more);
//  * other blocks (edges, other sizes, rejected descriptors) are predicted
//    one per wavefront, in turn.
// Every wavefront computes the same prefix sums and lists, so no barrier is
// needed.
__global__ __launch_bounds__(256) void mc_luma_kernel(const PicDev *__restrict__ table, int npics,
                                                      const tcme_pred *__restrict__ preds, int n,
                                                      uint8_t *__restrict__ out, int64_t out_bytes,
                                                      int32_t *__restrict__ status)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    constexpr int kXcd = 8;
    const int g = static_cast<int>(gridDim.x), bid = static_cast<int>(blockIdx.x);
    const int q = g / kXcd, rr = g % kXcd, x = bid % kXcd;
    const int chunk = x * q + (x < rr ? x : rr) + bid / kXcd;   // XCD x's chunks are contiguous
    const int64_t c0 = static_cast<int64_t>(n) * chunk / g;
    const int64_t c1 = static_cast<int64_t>(n) * (chunk + 1) / g;
    PicDev R{};
    int cached = -1;
    for (int64_t w0 = c0; w0 < c1; w0 += 64) {
        const int cnt = c1 - w0 < 64 ? static_cast<int>(c1 - w0) : 64;
        tcme_pred p{};
        bool ok = false;
        uint64_t an = 0;
        int apitch = 0, lg = 0;
        if (lane < cnt) {
            p = preds[w0 + lane];
            lg = p.blk == 8 ? 3 : p.blk == 16 ? 4 : 0;
            if (lg && p.pic >= 0 && p.pic < npics) {
                const PicDev Rk = table[p.pic];
                const int fx = p.mv_x >> 2, fy = p.mv_y >> 2;
                ok = Rk.y && Rk.an && span_fits(p.out_off, p.blk, p.blk, p.out_stride, out_bytes) &&
                     fx >= 2 && fy >= 2 && fx + p.blk + 3 <= Rk.w && fy + p.blk + 3 <= Rk.h;
                an = reinterpret_cast<uint64_t>(Rk.an);
                apitch = Rk.apitch;
            }
        }
        const int tasks = ok ? p.blk * p.blk / 8 : 0;
        int incl = tasks;  // inclusive prefix sum over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = shfl32(incl, lane >= d ? lane - d : lane);
            if (lane >= d) incl += t;
        }
        const int excl = incl - tasks;
        const int total = __builtin_amdgcn_readlane(incl, 63);
        if (status && ok && wave == 0) status[w0 + lane] = 0;
        for (int t0 = wave * 64; t0 < total; t0 += 256) {
            // lanes past the last task take the last task's block and rows
            // (so every address is one of an eligible block) and store nothing
            const int t = t0 + lane, tt = t < total ? t : total - 1;
            // the last descriptor whose tasks start at or before tt (excl is
            // non-decreasing; descriptors without tasks share the next one's
            // start and are passed over)
            int k = 0;
#pragma unroll
            for (int sft = 32; sft; sft >>= 1) {
                const int e = shfl32(excl, k + sft < 64 ? k + sft : 63);
                if (k + sft < cnt && e <= tt) k += sft;
            }
            const int u = tt - shfl32(excl, k), kl = shfl32(lg, k);
            const int mvx = shfl32(p.mv_x, k), mvy = shfl32(p.mv_y, k);
            const int os = shfl32(p.out_stride, k), ap = shfl32(apitch, k);
            const uint64_t a = shfl64(an, k), oo = shfl64(static_cast<uint64_t>(p.out_off), k);
            if (t < total)
                column8(reinterpret_cast<const uint64_t *>(a), ap, mvx, mvy, u & ((1 << kl) - 1),
                        (u >> kl) * 8, out + oo, os);
        }
        // 32x32 and 64x64 blocks: every wavefront a quarter of the rows
        uint64_t m = __ballot(lane < cnt && p.blk >= 32 && p.blk <= kMaxBlk);
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            const tcme_pred pk = preds[w0 + k];
            const int band = (pk.blk + 3) / 4;  // luma_rows clips the last band
            luma_rows(table, npics, pk, w0 + k, wave * band, (wave + 1) * band, out, out_bytes,
                      status, R, cached, lane);
        }
        // the rest (edges, other sizes, rejected descriptors), one
        // descriptor per wavefront in turn
        m = __ballot(lane < cnt && !ok && (p.blk < 32 || p.blk > kMaxBlk));
        for (int j = 0; m; ++j) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            if ((j & 3) == wave)
                luma_rows(table, npics, preds[w0 + k], w0 + k, 0, kMaxBlk, out, out_bytes, status,
                          R, cached, lane);
        }
    }
}

// Chroma sample (c, r) of a block at 1/8-sample MV (integer part xi, yi,
// fraction fx, fy): edge-clamped bilinear
// ((8-fx)(8-fy)A + fx(8-fy)B + (8-fx)fy C + fx fy D + 32) >> 6.
__device__ __forceinline__ int chroma_px(const PicDev &R, int xi, int yi, int fx, int fy, int c,
                                         int r)
{
    const gptr8 Y = (gptr8)R.y;
    const int x0 = min(max(xi + c, 0), R.w - 1), x1 = min(max(xi + c + 1, 0), R.w - 1);
    const int y0 = min(max(yi + r, 0), R.h - 1), y1 = min(max(yi + r + 1, 0), R.h - 1);
    const int a = Y[static_cast<size_t>(y0) * R.pitch + x0];
    const int b = Y[static_cast<size_t>(y0) * R.pitch + x1];
    const int cc = Y[static_cast<size_t>(y1) * R.pitch + x0];
    const int d = Y[static_cast<size_t>(y1) * R.pitch + x1];
    return ((8 - fx) * (8 - fy) * a + fx * (8 - fy) * b + (8 - fx) * fy * cc + fx * fy * d + 32) >> 6;
}

// Rows r0, r0 + dr, ... < n of column c of a chroma block, K rows per lane at
// a time, loads first; rows past the block re-read its last and are not stored.
template <int K>
__device__ __forceinline__ void chroma_rows(const PicDev &R, int xi, int yi, int fx, int fy,
                                            gout8 o, int ostride, int c, int r0, int dr, int n)
{
    for (int rb = r0; rb < n; rb += K * dr) {
        int v[K];
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = chroma_px(R, xi, yi, fx, fy, c, min(rb + k * dr, n - 1));
#pragma unroll
        for (int k = 0; k < K; ++k)
            if (rb + k * dr < n) o[static_cast<int64_t>(rb + k * dr) * ostride + c] = static_cast<uint8_t>(v[k]);
    }
}

// One chroma block by the whole wavefront, up to 4 rows per lane at a time
// (the descriptor index is wave-uniform: scalar descriptor and table loads).
__device__ __forceinline__ void chroma_block(const PicDev *__restrict__ table, int npics,
                                             const tcme_pred &p, int64_t i,
                                             uint8_t *__restrict__ out, int64_t out_bytes,
                                             int32_t *__restrict__ status, PicDev &R, int &cached,
                                             int lane)
{
    const int st = check_pred(table, npics, p, out_bytes, false, R, cached);
    if (status && lane == 0) status[i] = st;
    if (st) return;
    const int xi = p.mv_x >> 3, yi = p.mv_y >> 3, fx = p.mv_x & 7, fy = p.mv_y & 7;
    const gout8 o = (gout8)out + p.out_off;
    const BlockLanes bl(p.blk, lane);
    if ((p.blk + bl.dr - 1) / bl.dr >= 4)  // rows per lane; wave-uniform
        chroma_rows<4>(R, xi, yi, fx, fy, o, p.out_stride, bl.c, bl.r0, bl.dr, p.blk);
    else
        chroma_rows<1>(R, xi, yi, fx, fy, o, p.out_stride, bl.c, bl.r0, bl.dr, p.blk);
}

// Rows [r0, r0 + 4) of column c of a chroma block whose every sample's 2x2
// neighbourhood is inside the plane (no clamping), plane y / pitch, at
// 1/8-sample (mvx, mvy): the two samples (x, x+1) of the 5 rows are loaded
// first.
__device__ __forceinline__ void chroma_col4(const uint8_t *y, int pitch, int mvx, int mvy, int c,
                                            int r0, uint8_t *o, int os)
{
    const int xi = mvx >> 3, yi = mvy >> 3, fx = mvx & 7, fy = mvy & 7;
    const gptr8 p = (gptr8)y + static_cast<size_t>(yi + r0) * pitch + xi + c;
    int a[5], b[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        a[r] = p[static_cast<size_t>(r) * pitch];
        b[r] = p[static_cast<size_t>(r) * pitch + 1];
    }
    const int w00 = (8 - fx) * (8 - fy), w10 = fx * (8 - fy), w01 = (8 - fx) * fy, w11 = fx * fy;
    const gout8 oc = (gout8)o + static_cast<int64_t>(r0) * os + c;
#pragma unroll
    for (int r = 0; r < 4; ++r)
        oc[static_cast<int64_t>(r) * os] = static_cast<uint8_t>(
            (w00 * a[r] + w10 * b[r] + w01 * a[r + 1] + w11 * b[r + 1] + 32) >> 6);
}

// Chroma at 1/8 sample (tc_inter_predict_chroma, src/motion.c:597-632), the
// same work split as the luma kernel: a chunk of the list per workgroup,
// taken 64 descriptors at a time; blocks of 4x4 .. 32x32 whose samples need
// no edge clamping are cut into tasks of one column x 4 rows (m * m / 4 per
// block), numbered in a row over the 64 descriptors and taken by the
// workgroup's 256 lanes 256 at a time (chroma_col4: one round trip to memory
// per pass); other blocks are predicted one per wavefront, in turn.  (Quads
// of consecutive blocks per wavefront, four 4x4 blocks at 16 lanes each and
// larger ones one after another, took 63 us for a 1080p frame's 343,200
// blocks, most of it in the 16x16 and 32x32 blocks' rounds of loads.)
__global__ __launch_bounds__(256) void mc_chroma_kernel(const PicDev *__restrict__ table,
                                                        int npics,
                                                        const tcme_pred *__restrict__ preds,
                                                        int n, uint8_t *__restrict__ out,
                                                        int64_t out_bytes,
                                                        int32_t *__restrict__ status)
{
    const int lane = static_cast<int>(threadIdx.x & 63);
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    constexpr int kXcd = 8;
    const int g = static_cast<int>(gridDim.x), bid = static_cast<int>(blockIdx.x);
    const int q = g / kXcd, rr = g % kXcd, x = bid % kXcd;
    const int chunk = x * q + (x < rr ? x : rr) + bid / kXcd;   // XCD x's chunks are contiguous
    const int64_t c0 = static_cast<int64_t>(n) * chunk / g;
    const int64_t c1 = static_cast<int64_t>(n) * (chunk + 1) / g;
    PicDev R{};
    int cached = -1;
    for (int64_t w0 = c0; w0 < c1; w0 += 64) {
        const int cnt = c1 - w0 < 64 ? static_cast<int>(c1 - w0) : 64;
        tcme_pred p{};
        bool ok = false;
        uint64_t y = 0;
        int pitch = 0, lg = 0;
        if (lane < cnt) {
            p = preds[w0 + lane];
            lg = p.blk == 4 ? 2 : p.blk == 8 ? 3 : p.blk == 16 ? 4 : p.blk == 32 ? 5 : 0;
            if (lg && p.pic >= 0 && p.pic < npics) {
                const PicDev Rk = table[p.pic];
                const int xi = p.mv_x >> 3, yi = p.mv_y >> 3;
                ok = Rk.y && span_fits(p.out_off, p.blk, p.blk, p.out_stride, out_bytes) &&
                     xi >= 0 && yi >= 0 && xi + p.blk + 1 <= Rk.w && yi + p.blk + 1 <= Rk.h;
                y = reinterpret_cast<uint64_t>(Rk.y);
                pitch = Rk.pitch;
            }
        }
        const int tasks = ok ? p.blk * p.blk / 4 : 0;
        int incl = tasks;  // inclusive prefix sum over the lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = shfl32(incl, lane >= d ? lane - d : lane);
            if (lane >= d) incl += t;
        }
        const int excl = incl - tasks;
        const int total = __builtin_amdgcn_readlane(incl, 63);
        if (status && ok && wave == 0) status[w0 + lane] = 0;
        for (int t0 = wave * 64; t0 < total; t0 += 256) {
            const int t = t0 + lane, tt = t < total ? t : total - 1;
            int k = 0;  // the last descriptor whose tasks start at or before tt
#pragma unroll
            for (int sft = 32; sft; sft >>= 1) {
                const int e = shfl32(excl, k + sft < 64 ? k + sft : 63);
                if (k + sft < cnt && e <= tt) k += sft;
            }
            const int u = tt - shfl32(excl, k), kl = shfl32(lg, k);
            const int mvx = shfl32(p.mv_x, k), mvy = shfl32(p.mv_y, k);
            const int os = shfl32(p.out_stride, k), pt = shfl32(pitch, k);
            const uint64_t yy = shfl64(y, k), oo = shfl64(static_cast<uint64_t>(p.out_off), k);
            if (t < total)
                chroma_col4(reinterpret_cast<const uint8_t *>(yy), pt, mvx, mvy,
                            u & ((1 << kl) - 1), (u >> kl) * 4, out + oo, os);
        }
        uint64_t m = __ballot(lane < cnt && !ok);
        for (int j = 0; m; ++j) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            if ((j & 3) == wave)
                chroma_block(table, npics, preds[w0 + k], w0 + k, out, out_bytes, status, R,
                             cached, lane);
        }
    }
}

}  // namespace

hipError_t launch_predict(bool chroma, const PicDev *table, int npics, const tcme_pred *preds,
                          int n, uint8_t *out, int64_t out_bytes, int32_t *status, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    // a chunk of the list per workgroup (at least one descriptor)
    const dim3 grid(std::min(n, 8 * device_cus()));
    if (chroma)
        hipLaunchKernelGGL(mc_chroma_kernel, grid, dim3(256), 0, s, table, npics, preds, n, out,
                           out_bytes, status);
    else
        hipLaunchKernelGGL(mc_luma_kernel, grid, dim3(256), 0, s, table, npics, preds, n, out,
                           out_bytes, status);
    return hipGetLastError();
}

}  // namespace tcme
