// gfx950 kernels for TCodec's motion compensation (the consumers of each
// search result):
//
//   mc_luma_kernel    tc_inter_predict, reference src/motion.c:493-584, with
//                     get_interp_luma's per-sample ladder (src/motion.c:269-304)
//   mc_chroma_kernel  tc_inter_predict_chroma, src/motion.c:597-632
//
// One workgroup per block; a thread produces samples of one row.  Every
// branch of the reference's luma predictor (row copy for an in-frame integer
// MV, whole-block six-tap, per-sample six-tap / bilinear / nearest ladder)
// gives, sample for sample, the ladder's value, so the kernel computes the
// ladder directly: six-tap where the sample's 6x6 apron is inside the frame,
// bilinear where its 2x2 neighbourhood is, nearest (clamped) otherwise.  The
// six-tap value comes from the anchor plane (sample, hh, vv, cc per position)
// read as the half-sample lattice M[Y][X]: with X0 = q >> 1, X1 = (q+1) >> 1
// per axis at quarter-pel q,
//   v = clip((avg(M[Y0][X0], M[Y0][X1]) + avg(M[Y1][X0], M[Y1][X1]) + 1) >> 1),
// avg(a,b) = (a+b+1) >> 1 on raw values (src/motion.c:193-260; the h- and
// v-only paths of src/motion.c:133-191 agree after their final clip).
#include "tcme_internal.h"

namespace tcme {
namespace {

typedef const __attribute__((address_space(1))) uint8_t *gptr8;
typedef const __attribute__((address_space(1))) uint64_t *gptr64;
typedef __attribute__((address_space(1))) uint8_t *gout8;

__device__ __forceinline__ int clip255(int v) { return v < 0 ? 0 : (v > 255 ? 255 : v); }

// Lattice entry (Y, X): 16-bit field (Y&1)*2 + (X&1) of anchor (Y/2, X/2),
// sign-extended (the sample field is 0..255).
__device__ __forceinline__ int lattice(gptr64 an, int apitch, int Y, int X)
{
    const uint64_t a = an[static_cast<size_t>(Y >> 1) * apitch + (X >> 1)];
    const int sh = ((Y & 1) * 2 + (X & 1)) * 16;
    return static_cast<int>(static_cast<int16_t>(static_cast<uint16_t>(a >> sh)));
}

// Bilinear quarter-pel step, rounded per pass (src/motion.c:111-116).
__device__ __forceinline__ int lerp4(int a, int b, int f) { return (a * (4 - f) + b * f + 2) >> 2; }

// get_interp_luma (src/motion.c:269-304) at integer (x,y), fraction (dx,dy).
__device__ __forceinline__ int luma_px(const PicDev &R, int x, int y, int dx, int dy)
{
    const gptr8 Y = (gptr8)R.y;
    if (x >= 2 && x + 3 < R.w && y >= 2 && y + 3 < R.h) {
        if (dx == 0 && dy == 0) return Y[static_cast<size_t>(y) * R.pitch + x];
        const gptr64 an = (gptr64)R.an;
        const int qx = 4 * x + dx, qy = 4 * y + dy;
        const int x0 = qx >> 1, x1 = (qx + 1) >> 1, y0 = qy >> 1, y1 = (qy + 1) >> 1;
        const int top = (lattice(an, R.apitch, y0, x0) + lattice(an, R.apitch, y0, x1) + 1) >> 1;
        const int bot = (lattice(an, R.apitch, y1, x0) + lattice(an, R.apitch, y1, x1) + 1) >> 1;
        return clip255((top + bot + 1) >> 1);
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
__device__ __forceinline__ int check_pred(const PicDev *table, int npics, const tcme_pred &p,
                                          int64_t out_bytes, bool need_anchors, PicDev &R)
{
    if (p.pic < 0 || p.pic >= npics) return TCME_E_PIC;
    R = table[p.pic];
    if (!R.y || (need_anchors && !R.an)) return TCME_E_PIC;
    if (p.blk < 1 || p.blk > kMaxBlk) return TCME_E_BLK;
    if (p.out_stride < p.blk || p.out_off < 0 ||
        p.out_off + static_cast<int64_t>(p.blk - 1) * p.out_stride + p.blk > out_bytes)
        return TCME_E_ARG;
    return 0;
}

__global__ __launch_bounds__(256) void mc_luma_kernel(const PicDev *__restrict__ table, int npics,
                                                      const tcme_pred *__restrict__ preds, int n,
                                                      uint8_t *__restrict__ out, int64_t out_bytes,
                                                      int32_t *__restrict__ status)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const tcme_pred p = preds[i];
    PicDev R{};
    const int st = check_pred(table, npics, p, out_bytes, true, R);
    if (status && threadIdx.x == 0) status[i] = st;
    if (st) return;
    const int fx = p.mv_x >> 2, fy = p.mv_y >> 2, dx = p.mv_x & 3, dy = p.mv_y & 3;
    const gout8 o = (gout8)out + p.out_off;
    for (int k = threadIdx.x; k < p.blk * p.blk; k += blockDim.x) {
        const int r = k / p.blk, c = k - r * p.blk;
        o[static_cast<int64_t>(r) * p.out_stride + c] =
            static_cast<uint8_t>(luma_px(R, fx + c, fy + r, dx, dy));
    }
}

// Chroma at 1/8 sample: integer mv >> 3, fraction mv & 7, edge-clamped
// bilinear ((8-fx)(8-fy)A + fx(8-fy)B + (8-fx)fy C + fx fy D + 32) >> 6.
__global__ __launch_bounds__(256) void mc_chroma_kernel(const PicDev *__restrict__ table,
                                                        int npics,
                                                        const tcme_pred *__restrict__ preds,
                                                        int n, uint8_t *__restrict__ out,
                                                        int64_t out_bytes,
                                                        int32_t *__restrict__ status)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const tcme_pred p = preds[i];
    PicDev R{};
    const int st = check_pred(table, npics, p, out_bytes, false, R);
    if (status && threadIdx.x == 0) status[i] = st;
    if (st) return;
    const int xi = p.mv_x >> 3, yi = p.mv_y >> 3, fx = p.mv_x & 7, fy = p.mv_y & 7;
    const gptr8 Y = (gptr8)R.y;
    const gout8 o = (gout8)out + p.out_off;
    for (int k = threadIdx.x; k < p.blk * p.blk; k += blockDim.x) {
        const int r = k / p.blk, c = k - r * p.blk;
        const int y0 = min(max(yi + r, 0), R.h - 1), y1 = min(max(yi + r + 1, 0), R.h - 1);
        const int x0 = min(max(xi + c, 0), R.w - 1), x1 = min(max(xi + c + 1, 0), R.w - 1);
        const int a = Y[static_cast<size_t>(y0) * R.pitch + x0];
        const int b = Y[static_cast<size_t>(y0) * R.pitch + x1];
        const int cc = Y[static_cast<size_t>(y1) * R.pitch + x0];
        const int d = Y[static_cast<size_t>(y1) * R.pitch + x1];
        const int v = (8 - fx) * (8 - fy) * a + fx * (8 - fy) * b + (8 - fx) * fy * cc + fx * fy * d;
        o[static_cast<int64_t>(r) * p.out_stride + c] = static_cast<uint8_t>((v + 32) >> 6);
    }
}

}  // namespace

hipError_t launch_predict(bool chroma, const PicDev *table, int npics, const tcme_pred *preds,
                          int n, uint8_t *out, int64_t out_bytes, int32_t *status, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (chroma)
        hipLaunchKernelGGL(mc_chroma_kernel, dim3(n), dim3(256), 0, s, table, npics, preds, n, out,
                           out_bytes, status);
    else
        hipLaunchKernelGGL(mc_luma_kernel, dim3(n), dim3(256), 0, s, table, npics, preds, n, out,
                           out_bytes, status);
    return hipGetLastError();
}

}  // namespace tcme
