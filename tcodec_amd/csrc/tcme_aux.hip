// gfx950 kernels for the encoder-side helpers around the search:
//
//   global_mv_kernel  estimate_global_mv, reference src/encoder.c:171-202: a
//                     greedy two-pass walk over whole-frame coarse SADs (the
//                     ME centre hint used when |glob| > 16 quarter-pel,
//                     src/encoder.c:1000-1004)
//   sad_kernel        block SADs: the fast-mode SAD proxies of a prediction
//                     (src/encoder.c:814, 839, 955, 1171) and the co-located
//                     CTU / 8x8 stability SADs (src/encoder.c:2843-2891)
#include "tcme_internal.h"

namespace tcme {
namespace {

typedef const __attribute__((address_space(1))) uint8_t *gptr8;

constexpr int kGmvThreads = 512;

__device__ __forceinline__ int64_t block_reduce(int v, int *red)
{
    // wave sums with DPP row scans + readlanes, then across wavefronts in LDS
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);
    const int w = __builtin_amdgcn_readlane(v, 15) + __builtin_amdgcn_readlane(v, 31) +
                  __builtin_amdgcn_readlane(v, 47) + __builtin_amdgcn_readlane(v, 63);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) red[wave] = w;
    __syncthreads();
    // every thread sums the same LDS values: the total is uniform
    int64_t t = 0;
#pragma unroll
    for (int k = 0; k < kGmvThreads / 64; ++k) t += red[k];
    __syncthreads();  // red is reused by the next reduction
    return t;
}

// One workgroup per (current, reference) pair.  Samples sit every 8 pixels
// (every other position of the 1/4-resolution grid, src/encoder.c:184-185);
// a sample displaced off the frame costs 255*4 (src/encoder.c:190).  The
// walk is the reference's: candidates in row-major order around the moving
// best, the first of each pass always replacing the initial 2^60 bound.
__global__ __launch_bounds__(kGmvThreads) void global_mv_kernel(const PicDev *__restrict__ table,
                                                                int npics,
                                                                const int32_t *__restrict__ pairs,
                                                                int n, int32_t *__restrict__ out)
{
    __shared__ int red[kGmvThreads / 64];
    const int i = blockIdx.x;
    if (i >= n) return;
    const int cp = pairs[2 * i], rp = pairs[2 * i + 1];
    int status = 0;
    PicDev C{}, R{};
    if (cp < 0 || cp >= npics || rp < 0 || rp >= npics) status = TCME_E_PIC;
    else {
        C = table[cp];
        R = table[rp];
        if (!C.y || !R.y) status = TCME_E_PIC;
        else if (C.w != R.w || C.h != R.h) status = TCME_E_ARG;
    }
    if (status) {
        if (threadIdx.x == 0) { out[3 * i] = 0; out[3 * i + 1] = 0; out[3 * i + 2] = status; }
        return;
    }
    const int w = C.w, h = C.h, cw = w / 4, ch = h / 4;
    const int nx = (cw + 1) / 2, ny = (ch + 1) / 2;   // x, y = 0, 2, 4, .. < cw, ch
    const int ns = nx * ny;
    const gptr8 cy = (gptr8)C.y, ry = (gptr8)R.y;
    int bx = 0, by = 0;
    int64_t best = static_cast<int64_t>(1) << 60;
    for (int pass = 0; pass < 2; ++pass) {
        const int step = pass ? 1 : 2, range = pass ? 1 : 8;
        bool improved = false;
        for (int dy = -range; dy <= range; dy += step)
            for (int dx = -range; dx <= range; dx += step) {
                int s = 0;
                for (int k = threadIdx.x; k < ns; k += kGmvThreads) {
                    const int y = 2 * (k / nx), x = 2 * (k - (k / nx) * nx);
                    const int rx = x * 4 + bx + dx * 4, ry_ = y * 4 + by + dy * 4;
                    if (rx < 0 || ry_ < 0 || rx >= w || ry_ >= h) { s += 255 * 4; continue; }
                    const int o = cy[static_cast<size_t>(y * 4) * C.pitch + x * 4];
                    const int r = ry[static_cast<size_t>(ry_) * R.pitch + rx];
                    s += o > r ? o - r : r - o;
                }
                const int64_t tot = block_reduce(s, red);
                if (tot < best) {
                    best = tot;
                    bx += dx * 4;
                    by += dy * 4;
                    improved = true;
                }
            }
        if (!improved && pass) break;
    }
    if (threadIdx.x == 0) {
        out[3 * i] = bx * 4;
        out[3 * i + 1] = by * 4;
        out[3 * i + 2] = 0;
    }
}

// One wavefront per descriptor; lanes take samples, a DPP scan sums them.
__global__ __launch_bounds__(256) void sad_kernel(const PicDev *__restrict__ table, int npics,
                                                  const tcme_sad_desc *__restrict__ descs, int n,
                                                  const uint8_t *__restrict__ buf,
                                                  int64_t buf_bytes, int32_t *__restrict__ out)
{
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= n) return;  // a whole wavefront leaves together
    const tcme_sad_desc d = descs[i];
    int status = 0;
    PicDev A{}, B{};
    if (d.pic < 0 || d.pic >= npics || d.other_pic >= npics) status = TCME_E_PIC;
    else if (d.w < 1 || d.h < 1 || d.w > kMaxBlk || d.h > kMaxBlk) status = TCME_E_BLK;
    else {
        A = table[d.pic];
        if (!A.y) status = TCME_E_PIC;
        else if (d.x < 0 || d.y < 0 || d.x + d.w > A.w || d.y + d.h > A.h) status = TCME_E_ARG;
        else if (d.other_pic >= 0) {
            B = table[d.other_pic];
            if (!B.y) status = TCME_E_PIC;
            else if (d.ox < 0 || d.oy < 0 || d.ox + d.w > B.w || d.oy + d.h > B.h)
                status = TCME_E_ARG;
        } else if (!buf || d.buf_stride < d.w || d.buf_off < 0 ||
                   d.buf_off + static_cast<int64_t>(d.h - 1) * d.buf_stride + d.w > buf_bytes) {
            status = TCME_E_ARG;
        }
    }
    if (status) {
        if (lane == 0) out[i] = status;
        return;
    }
    const gptr8 a = (gptr8)A.y + static_cast<size_t>(d.y) * A.pitch + d.x;
    const gptr8 b = d.other_pic >= 0 ? (gptr8)B.y + static_cast<size_t>(d.oy) * B.pitch + d.ox
                                     : (gptr8)buf + d.buf_off;
    const int bp = d.other_pic >= 0 ? B.pitch : d.buf_stride;
    int s = 0;
    for (int k = lane; k < d.w * d.h; k += 64) {
        const int r = k / d.w, c = k - r * d.w;
        const int p = a[static_cast<size_t>(r) * A.pitch + c], q = b[static_cast<size_t>(r) * bp + c];
        s += p > q ? p - q : q - p;
    }
    s += __builtin_amdgcn_update_dpp(0, s, 0x111, 0xF, 0xF, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x112, 0xF, 0xF, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xF, false);
    s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xF, 0xF, false);
    const int t = __builtin_amdgcn_readlane(s, 15) + __builtin_amdgcn_readlane(s, 31) +
                  __builtin_amdgcn_readlane(s, 47) + __builtin_amdgcn_readlane(s, 63);
    if (lane == 0) out[i] = t;
}

}  // namespace

hipError_t launch_global_mv(const PicDev *table, int npics, const int32_t *pairs, int n,
                            int32_t *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(global_mv_kernel, dim3(n), dim3(kGmvThreads), 0, s, table, npics, pairs, n,
                       out);
    return hipGetLastError();
}

hipError_t launch_sad(const PicDev *table, int npics, const tcme_sad_desc *descs, int n,
                      const uint8_t *buf, int64_t buf_bytes, int32_t *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sad_kernel, dim3((n + 3) / 4), dim3(256), 0, s, table, npics, descs, n, buf,
                       buf_bytes, out);
    return hipGetLastError();
}

}  // namespace tcme
