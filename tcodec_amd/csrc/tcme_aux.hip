// gfx950 kernels for the encoder-side helpers around the search:
//
//   gmv_table_kernel  estimate_global_mv, reference src/encoder.c:171-202: a
//   gmv_walk_kernel   greedy two-pass walk over whole-frame coarse SADs (the
//                     ME centre hint used when |glob| > 16 quarter-pel,
//                     src/encoder.c:1000-1004): a parallel table of the SADs
//                     of every displacement within reach, then the walk
//   sad_kernel        block SADs: the fast-mode SAD proxies of a prediction
//                     (src/encoder.c:814, 839, 955, 1171) and the co-located
//                     CTU / 8x8 stability SADs (src/encoder.c:2843-2891)
#include <algorithm>

#include "tcme_internal.h"

namespace tcme {
namespace {

typedef const __attribute__((address_space(1))) uint8_t *gptr8;

constexpr int kGmvThreads = 512;
constexpr int kGmvReach = 128;                    // table covers |Dx|, |Dy| <= 128 px
constexpr int kGmvSide = 2 * kGmvReach / 4 + 1;   // 65 displacements per axis
constexpr int kGmvEntries = kGmvSide * kGmvSide;  // 4225
constexpr int kGmvLdsMax = 48 * 1024;             // table kernel's staging bound

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

// Coarse SAD of one candidate displacement (Dx, Dy) px, all threads of the
// workgroup: samples every 8 pixels (every other position of the 1/4-resolution
// grid, src/encoder.c:184-185); a sample displaced off the frame costs 255*4
// (src/encoder.c:190).  The result is uniform.
__device__ int64_t gmv_direct(const PicDev &C, const PicDev &R, int nx, int ns, int Dx, int Dy,
                              int *red)
{
    const gptr8 cy = (gptr8)C.y, ry = (gptr8)R.y;
    int s = 0;
    for (int k = threadIdx.x; k < ns; k += kGmvThreads) {
        const int j = k / nx, i = k - j * nx;
        const int rx = 8 * i + Dx, ry_ = 8 * j + Dy;
        if (rx < 0 || ry_ < 0 || rx >= C.w || ry_ >= C.h) { s += 255 * 4; continue; }
        const int o = cy[static_cast<size_t>(8 * j) * C.pitch + 8 * i];
        const int r = ry[static_cast<size_t>(ry_) * R.pitch + rx];
        s += o > r ? o - r : r - o;
    }
    return block_reduce(s, red);
}

__device__ __forceinline__ int gmv_check(const PicDev *table, int npics, int cp, int rp,
                                         PicDev &C, PicDev &R)
{
    if (cp < 0 || cp >= npics || rp < 0 || rp >= npics) return TCME_E_PIC;
    C = table[cp];
    R = table[rp];
    if (!C.y || !R.y) return TCME_E_PIC;
    if (C.w != R.w || C.h != R.h) return TCME_E_ARG;
    return 0;
}

// Every displacement the walk can visit is a multiple of 4 px (pass 0 moves
// by 8, pass 1 by 4, from 0), so the coarse SADs of all displacements within
// +-kGmvReach px are computed up front, in parallel: entry (iy, ix) holds
// (Dx, Dy) = (4 ix - kGmvReach, 4 iy - kGmvReach).  Workgroup (chunk, band,
// pair) takes 256 entries and the band's `rows` sample rows j (picture row
// 8j), stages the band's current samples and the reference rows/columns at
// 4-px spacing those entries read, and writes one partial sum per entry.
// A thread owns one entry.  Entry ix reads staged column 2i + ix for sample
// i, so each staged reference row is kept as two halves (even columns, odd
// columns): the reads of consecutive samples are then consecutive bytes of
// one half, four samples per v_sad_u8.
__host__ __device__ constexpr int gmv_half(int nx) { return ((nx + kGmvSide / 2 + 1 + 3) & ~3) + 4; }
__host__ __device__ constexpr int gmv_curp(int nx) { return (nx + 3) & ~3; }

__global__ __launch_bounds__(256) void gmv_table_kernel(const PicDev *__restrict__ table,
                                                        int npics,
                                                        const int32_t *__restrict__ pairs,
                                                        int rows, int nbands,
                                                        int32_t *__restrict__ part)
{
    extern __shared__ uint32_t gmv_lds[];
    const int pair = blockIdx.z, band = blockIdx.y;
    PicDev C{}, R{};
    if (gmv_check(table, npics, pairs[2 * pair], pairs[2 * pair + 1], C, R)) return;
    const int w = C.w, h = C.h, nx = (w / 4 + 1) / 2, ny = (h / 4 + 1) / 2;
    const int j0 = band * rows;
    if (j0 >= ny) return;
    const int nj = min(rows, ny - j0);
    const int e0 = blockIdx.x * 256, e = e0 + threadIdx.x;
    const int iy_lo = e0 / kGmvSide, iy_hi = min(e0 + 255, kGmvEntries - 1) / kGmvSide;
    // staged reference rows: picture row 4*r4, r4 = 2j + iy - kGmvReach/4;
    // staged column k = c4 + q (picture column 4*c4), k = 0 .. lc-1, row r's
    // half p (= k & 1) at byte (2r + p) * hp + (k >> 1)
    const int q = kGmvReach / 4;
    const int r4_base = 2 * j0 + iy_lo - q;
    const int nr = 2 * (nj - 1) + (iy_hi - iy_lo) + 1;
    const int lc = 2 * nx + 2 * q + 1;
    const int hp = gmv_half(nx), cp = gmv_curp(nx);
    uint8_t *ref_s = reinterpret_cast<uint8_t *>(gmv_lds);
    uint8_t *cur_s = ref_s + 2 * nr * hp;
    const gptr8 cy = (gptr8)C.y, ry = (gptr8)R.y;
    for (int k = threadIdx.x; k < nr * 2 * hp; k += 256) {
        const int rp = k / hp, m = k - rp * hp;
        const int r = rp >> 1, c = 2 * m + (rp & 1);
        const int c4 = c - q, r4 = r4_base + r;
        const bool in = c < lc && r4 >= 0 && 4 * r4 < h && c4 >= 0 && 4 * c4 < w;
        ref_s[k] = in ? ry[static_cast<size_t>(4 * r4) * R.pitch + 4 * c4] : 0;
    }
    for (int k = threadIdx.x; k < nj * cp; k += 256) {
        const int jj = k / cp, i = k - jj * cp;
        cur_s[k] = i < nx ? cy[static_cast<size_t>(8 * (j0 + jj)) * C.pitch + 8 * i] : 0;
    }
    __syncthreads();
    if (e >= kGmvEntries) return;
    const int iy = e / kGmvSide, ix = e - iy * kGmvSide;
    const int cw4 = (w + 3) / 4;  // c4 in frame: 0 <= c4 < cw4
    // samples i whose column 2i + ix - q is in frame: [ilo, ihi]
    const int ilo = max(0, (q - ix + 1) >> 1);
    const int ihi = min(nx - 1, (cw4 - 1 + q - ix) >> 1);
    const int nin = ihi >= ilo ? ihi - ilo + 1 : 0;
    const int ia = (ilo + 3) & ~3, ib = (ihi + 1) & ~3;  // whole dwords of samples
    const int hofs = ix >> 1;
    const uint32_t sh = static_cast<uint32_t>(hofs & 3);
    int acc = 0;
    for (int jj = 0; jj < nj; ++jj) {
        const int r4 = 2 * (j0 + jj) + iy - q;
        if (r4 < 0 || 4 * r4 >= h) { acc += nx * (255 * 4); continue; }
        acc += (nx - nin) * (255 * 4);
        const uint8_t *rr = ref_s + (2 * (r4 - r4_base) + (ix & 1)) * hp + hofs;  // rr[i]
        const uint8_t *cr = cur_s + jj * cp;
        if (ia >= ib) {
            for (int i = ilo; i <= ihi; ++i) acc += abs(static_cast<int>(cr[i]) - static_cast<int>(rr[i]));
            continue;
        }
        for (int i = ilo; i < ia; ++i) acc += abs(static_cast<int>(cr[i]) - static_cast<int>(rr[i]));
        const uint32_t *rw = reinterpret_cast<const uint32_t *>(ref_s + (2 * (r4 - r4_base) + (ix & 1)) * hp);
        const uint32_t *cw = reinterpret_cast<const uint32_t *>(cr);
        uint32_t vacc = 0;
#pragma unroll 8
        for (int i = ia; i < ib; i += 4) {
            const int b = (i + hofs) >> 2;
            vacc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(rw[b + 1], rw[b], sh), cw[i >> 2],
                                           vacc);
        }
        acc += static_cast<int>(vacc);
        for (int i = ib; i <= ihi; ++i) acc += abs(static_cast<int>(cr[i]) - static_cast<int>(rr[i]));
    }
    part[(static_cast<size_t>(pair) * nbands + band) * kGmvEntries + e] = acc;
}

// One workgroup per (current, reference) pair: sums the bands of the table,
// then walks as the reference does, candidates in row-major order around the
// moving best, the first of each pass always replacing the initial 2^60
// bound (src/encoder.c:178-197).  A displacement outside the table is priced
// directly.
__global__ __launch_bounds__(kGmvThreads) void gmv_walk_kernel(const PicDev *__restrict__ table,
                                                               int npics,
                                                               const int32_t *__restrict__ pairs,
                                                               int n, int rows, int nbands,
                                                               const int32_t *__restrict__ part,
                                                               int32_t *__restrict__ out)
{
    __shared__ int red[kGmvThreads / 64];
    __shared__ int64_t tab[kGmvEntries];
    const int i = blockIdx.x;
    if (i >= n) return;
    PicDev C{}, R{};
    const int status = gmv_check(table, npics, pairs[2 * i], pairs[2 * i + 1], C, R);
    if (status) {
        if (threadIdx.x == 0) { out[3 * i] = 0; out[3 * i + 1] = 0; out[3 * i + 2] = status; }
        return;
    }
    const int nx = (C.w / 4 + 1) / 2, ny = (C.h / 4 + 1) / 2, ns = nx * ny;
    const int nb = (ny + rows - 1) / rows;
    // the table is the sum of the bands' partial tables: bands outside, this
    // thread's entries inside, so each band's loads are in flight together
    // (entries outside made every load wait for the one before: 81 in a row)
    constexpr int KE = (kGmvEntries + kGmvThreads - 1) / kGmvThreads;
    int64_t acc[KE] = {};
    for (int b = 0; b < nb; ++b) {
        const int32_t *q = part + (static_cast<size_t>(i) * nbands + b) * kGmvEntries;
#pragma unroll
        for (int k = 0; k < KE; ++k) {
            const int e = static_cast<int>(threadIdx.x) + k * kGmvThreads;
            if (e < kGmvEntries) acc[k] += q[e];
        }
    }
#pragma unroll
    for (int k = 0; k < KE; ++k) {
        const int e = static_cast<int>(threadIdx.x) + k * kGmvThreads;
        if (e < kGmvEntries) tab[e] = acc[k];
    }
    __syncthreads();
    int bx = 0, by = 0;
    int64_t best = static_cast<int64_t>(1) << 60;
    for (int pass = 0; pass < 2; ++pass) {
        const int step = pass ? 1 : 2, range = pass ? 1 : 8;
        bool improved = false;
        for (int dy = -range; dy <= range; dy += step)
            for (int dx = -range; dx <= range; dx += step) {
                const int Dx = bx + dx * 4, Dy = by + dy * 4;
                const int64_t tot =
                    (Dx >= -kGmvReach && Dx <= kGmvReach && Dy >= -kGmvReach && Dy <= kGmvReach)
                        ? tab[((Dy + kGmvReach) >> 2) * kGmvSide + ((Dx + kGmvReach) >> 2)]
                        : gmv_direct(C, R, nx, ns, Dx, Dy, red);
                if (tot < best) {
                    best = tot;
                    bx = Dx;
                    by = Dy;
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

// One wavefront per descriptor, the wavefronts striding over the list; lanes
// take samples, a DPP scan sums them.
// |a - b| summed over rows r0, r0 + dr, ... < h of one column, K rows per lane
// at a time with every load issued first (rows past the block re-read its
// last row and are not counted).
template <int K>
__device__ __forceinline__ int sad_rows(gptr8 a, int ap, gptr8 b, int bp, int r0, int dr, int h)
{
    int s = 0;
    for (int rb = r0; rb < h; rb += K * dr) {
        int pa[K], pb[K];
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int r = min(rb + k * dr, h - 1);
            pa[k] = a[static_cast<size_t>(r) * ap];
            pb[k] = b[static_cast<size_t>(r) * bp];
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int d = pa[k] - pb[k];
            if (rb + k * dr < h) s += d < 0 ? -d : d;
        }
    }
    return s;
}

// One wavefront per block, the wavefronts striding over the list; the block
// index is wave-uniform, so descriptor and picture-table entries are scalar
// loads.  Lane l takes column l mod w of rows l / w, l / w + 64 / w, ...
// (lanes past the last whole row idle), up to 8 rows at a time.
__global__ __launch_bounds__(256) void sad_kernel(const PicDev *__restrict__ table, int npics,
                                                  const tcme_sad_desc *__restrict__ descs, int n,
                                                  const uint8_t *__restrict__ buf,
                                                  int64_t buf_bytes, int32_t *__restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));
    const int stride = static_cast<int>(gridDim.x) * 4;
    for (int i = static_cast<int>(blockIdx.x) * 4 + wave; i < n; i += stride) {
        const tcme_sad_desc d = descs[i];
        int status = 0;
        PicDev A{}, B{};
        if (d.pic < 0 || d.pic >= npics || d.other_pic >= npics) status = TCME_E_PIC;
        else if (d.w < 1 || d.h < 1 || d.w > kMaxBlk || d.h > kMaxBlk) status = TCME_E_BLK;
        else {
            A = table[d.pic];
            if (!A.y) status = TCME_E_PIC;
            else if (!rect_fits(d.x, d.y, d.w, d.h, A.w, A.h)) status = TCME_E_ARG;
            else if (d.other_pic >= 0) {
                B = table[d.other_pic];
                if (!B.y) status = TCME_E_PIC;
                else if (!rect_fits(d.ox, d.oy, d.w, d.h, B.w, B.h))
                    status = TCME_E_ARG;
            } else if (!buf || !span_fits(d.buf_off, d.w, d.h, d.buf_stride, buf_bytes)) {
                status = TCME_E_ARG;
            }
        }
        if (status) {
            if (lane == 0) out[i] = status;
            continue;
        }
        const gptr8 a = (gptr8)A.y + static_cast<size_t>(d.y) * A.pitch + d.x;
        const gptr8 b = d.other_pic >= 0 ? (gptr8)B.y + static_cast<size_t>(d.oy) * B.pitch + d.ox
                                         : (gptr8)buf + d.buf_off;
        const int bp = d.other_pic >= 0 ? B.pitch : d.buf_stride;
        const int dr = 64 / d.w, c = lane % d.w, r0 = lane / d.w;
        const int rows = (d.h + dr - 1) / dr;  // per lane; wave-uniform
        int s = 0;
        if (r0 < dr) {
            if (rows >= 8) s = sad_rows<8>(a + c, A.pitch, b + c, bp, r0, dr, d.h);
            else if (rows >= 4) s = sad_rows<4>(a + c, A.pitch, b + c, bp, r0, dr, d.h);
            else s = sad_rows<1>(a + c, A.pitch, b + c, bp, r0, dr, d.h);
        }
        s += __builtin_amdgcn_update_dpp(0, s, 0x111, 0xF, 0xF, false);
        s += __builtin_amdgcn_update_dpp(0, s, 0x112, 0xF, 0xF, false);
        s += __builtin_amdgcn_update_dpp(0, s, 0x114, 0xF, 0xF, false);
        s += __builtin_amdgcn_update_dpp(0, s, 0x118, 0xF, 0xF, false);
        const int t = __builtin_amdgcn_readlane(s, 15) + __builtin_amdgcn_readlane(s, 31) +
                      __builtin_amdgcn_readlane(s, 47) + __builtin_amdgcn_readlane(s, 63);
        if (lane == 0) out[i] = t;
    }
}

}  // namespace

// Staging bytes of a table workgroup: both halves of up to 2(rows-1)+5
// reference rows, `rows` current sample rows.
size_t gmv_lds(int rows, int nx)
{
    const int nr = 2 * (rows - 1) + 5;
    return static_cast<size_t>(2 * nr) * gmv_half(nx) + static_cast<size_t>(rows) * gmv_curp(nx);
}

int gmv_rows(int max_w)
{
    const int nx = (max_w / 4 + 1) / 2;
    for (int rows = 16; rows > 1; rows >>= 1)
        if (gmv_lds(rows, nx) <= static_cast<size_t>(kGmvLdsMax)) return rows;
    return 1;
}

size_t gmv_part_bytes(int n, int max_w, int max_h)
{
    const int rows = gmv_rows(max_w), ny = (max_h / 4 + 1) / 2;
    const int nbands = std::max(1, (ny + rows - 1) / rows);
    return static_cast<size_t>(n) * nbands * kGmvEntries * sizeof(int32_t);
}

hipError_t launch_global_mv(const PicDev *table, int npics, const int32_t *pairs, int n,
                            int max_w, int max_h, int32_t *part, int32_t *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int rows = gmv_rows(max_w), nx = (max_w / 4 + 1) / 2, ny = (max_h / 4 + 1) / 2;
    const int nbands = std::max(1, (ny + rows - 1) / rows);
    const size_t lds = gmv_lds(rows, nx);
    const int chunks = (kGmvEntries + 255) / 256;
    hipLaunchKernelGGL(gmv_table_kernel, dim3(chunks, nbands, n), dim3(256), lds, s, table, npics,
                       pairs, rows, nbands, part);
    if (hipError_t e = hipGetLastError()) return e;
    hipLaunchKernelGGL(gmv_walk_kernel, dim3(n), dim3(kGmvThreads), 0, s, table, npics, pairs, n,
                       rows, nbands, static_cast<const int32_t *>(part), out);
    return hipGetLastError();
}

namespace {

// One call list for a tick's pictures: call i belongs to the picture p whose
// range [first, first + ncalls) holds it (binary search over the pictures'
// starts); it is that picture's layout entry i - first with the reference
// index mapped to the picture id and the current picture filled in.
__global__ __launch_bounds__(256) void expand_kernel(const TickPicDev *__restrict__ pics, int npics,
                                                     tcme_call *__restrict__ out, int64_t total)
{
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x; i < total;
         i += static_cast<int64_t>(gridDim.x) * 256) {
        int lo = 0, hi = npics - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (pics[mid].first <= i) lo = mid; else hi = mid - 1;
        }
        const TickPicDev &p = pics[lo];
        // the template pointer was read from memory: loaded through a
        // global view (two 16-byte words), so the load is not a generic one
        typedef unsigned int w4 __attribute__((ext_vector_type(4)));
        const __attribute__((address_space(1))) w4 *q =
            (const __attribute__((address_space(1))) w4 *)(p.tmpl) + 2 * (i - p.first);
        const w4 a = q[0], b = q[1];
        tcme_call c{static_cast<int32_t>(a.x), static_cast<int32_t>(a.y), static_cast<int32_t>(a.z),
                    static_cast<int32_t>(a.w), static_cast<int32_t>(b.x), static_cast<int32_t>(b.y),
                    static_cast<int32_t>(b.z), static_cast<int32_t>(b.w)};
        c.ref_pic = p.ref_pic[c.ref_pic];
        c.cur_pic = p.cur_pic;
        out[i] = c;
    }
}

}  // namespace

hipError_t launch_expand(const TickPicDev *pics, int npics, tcme_call *out, int64_t total,
                         hipStream_t s)
{
    if (total <= 0) return hipSuccess;
    const int64_t blocks = std::min<int64_t>((total + 255) / 256, 16 * device_cus());
    hipLaunchKernelGGL(expand_kernel, dim3(static_cast<unsigned>(blocks)), dim3(256), 0, s, pics,
                       npics, out, total);
    return hipGetLastError();
}

hipError_t launch_sad(const PicDev *table, int npics, const tcme_sad_desc *descs, int n,
                      const uint8_t *buf, int64_t buf_bytes, int32_t *out, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(sad_kernel, dim3(std::min((n + 3) / 4, 8 * device_cus())), dim3(256), 0, s,
                       table, npics, descs, n, buf, buf_bytes, out);
    return hipGetLastError();
}

}  // namespace tcme
