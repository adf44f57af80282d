// gfx950 kernels for TCodec's motion-estimation path.
//
//   halfpel_kernel    raw six-tap half-sample planes of a reference picture
//   me_search_kernel  the hierarchical hexagon search + quarter-pel refinement,
//                     one wavefront per call (reference src/motion.c:384-485)
//
// The search is the reference's sequential greedy walk, made wave-parallel
// without changing its answer: every candidate of a pattern pass is priced at
// once around the current best, the FIRST candidate (table order) that is
// strictly better is taken, and the candidates after it are priced again
// around the new best.  That is exactly what the sequential loop computes,
// because the candidates before the first improvement saw the same centre.
//
// Block costs use v_sad_u8 on packed bytes (4 |a-b| per lane-op) and DPP
// row reductions; the walk's control flow is wave-uniform (scalar unit).
#include "tcme_internal.h"

namespace tcme {
namespace {

constexpr int32_t kInf = 0x7FFFFFFF;

__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f)
{
    return (a + f - 5 * (b + e) + 20 * (c + d) + 16) >> 5;
}

// Global-memory view of a plane (address space 1), so plane reads compile to
// global loads rather than flat ones.
typedef const __attribute__((address_space(1))) uint32_t *gptr32;

__device__ __forceinline__ gptr32 as_global32(const void *p)
{
    return (gptr32)(p);
}

// Four bytes at byte offset `off` of a 4-byte-aligned buffer, assembled from
// the aligned dword pair that covers them.  Offsets, never integer-cast
// pointers, so every access keeps its address space.
__device__ __forceinline__ uint32_t ld4g(gptr32 base, size_t off)
{
    gptr32 q = base + (off >> 2);
    return __builtin_amdgcn_alignbyte(q[1], q[0], static_cast<uint32_t>(off & 3));
}

// The LDS reference window of the chunked kernel (dynamic shared memory).
extern __shared__ __align__(16) uint32_t tcme_win[];

__device__ __forceinline__ uint32_t ld4s(int off)
{
    const uint32_t *q = tcme_win + (off >> 2);
    return __builtin_amdgcn_alignbyte(q[1], q[0], static_cast<uint32_t>(off & 3));
}

// Inclusive sum scan inside each 16-lane DPP row; lane 16r+15 ends with the
// row total.  All 64 lanes must be active.
__device__ __forceinline__ int row_scan(int v)
{
    v += __builtin_amdgcn_update_dpp(0, v, 0x111, 0xF, 0xF, false);  // row_shr:1
    v += __builtin_amdgcn_update_dpp(0, v, 0x112, 0xF, 0xF, false);  // row_shr:2
    v += __builtin_amdgcn_update_dpp(0, v, 0x114, 0xF, 0xF, false);  // row_shr:4
    v += __builtin_amdgcn_update_dpp(0, v, 0x118, 0xF, 0xF, false);  // row_shr:8
    return v;
}

__device__ __forceinline__ int wave_sum(int v)
{
    v = row_scan(v);
    return __builtin_amdgcn_readlane(v, 15) + __builtin_amdgcn_readlane(v, 31) +
           __builtin_amdgcn_readlane(v, 47) + __builtin_amdgcn_readlane(v, 63);
}

// Search patterns after the implicit {0,0} (reference src/motion.c:347-366).
enum { PAT_LARGE, PAT_SMALL, PAT_DIAMOND, PAT_QPEL };
template <int P> struct Pat;
template <> struct Pat<PAT_LARGE> {
    static constexpr int K = 6;
    __device__ static constexpr int x(int k) { return k == 0 ? -2 : k == 1 ? 2 : k == 2 ? -4 : k == 3 ? 4 : k == 4 ? -2 : 2; }
    __device__ static constexpr int y(int k) { return k < 2 ? -4 : k < 4 ? 0 : 4; }
};
template <> struct Pat<PAT_SMALL> {
    static constexpr int K = 6;
    __device__ static constexpr int x(int k) { return k == 0 ? -1 : k == 1 ? 1 : k == 2 ? -2 : k == 3 ? 2 : k == 4 ? -1 : 1; }
    __device__ static constexpr int y(int k) { return k < 2 ? -2 : k < 4 ? 0 : 2; }
};
template <> struct Pat<PAT_DIAMOND> {
    static constexpr int K = 4;
    __device__ static constexpr int x(int k) { return k == 1 ? -1 : k == 2 ? 1 : 0; }
    __device__ static constexpr int y(int k) { return k == 0 ? -1 : k == 3 ? 1 : 0; }
};
template <> struct Pat<PAT_QPEL> {
    static constexpr int K = 8;
    __device__ static constexpr int x(int k) { return (k == 0 || k == 4 || k == 6) ? 1 : (k == 2 || k == 3) ? 0 : -1; }
    __device__ static constexpr int y(int k) { return (k == 0 || k == 1) ? 0 : (k == 2 || k == 4 || k == 5) ? 1 : -1; }
};

// Lane geometry for an N x N block held as dword units (4 pixels each).
// N >= 16: every lane owns U units of one candidate (G == 1).
// N == 8 : 16 units, so the wave's four DPP rows price four candidates (G == 4).
template <int N> struct Geo {
    static constexpr int DW = N / 4;
    static constexpr int UNITS = N * DW;
    static constexpr int U = UNITS >= 64 ? UNITS / 64 : 1;
    static constexpr int G = UNITS >= 64 ? 1 : 64 / UNITS;
};

struct Counts { int ie = 0, oob = 0, q6 = 0, qb = 0; };

__device__ __forceinline__ bool fits(const PicDev &R, int n, int x, int y)
{
    return x >= 0 && y >= 0 && x + n <= R.w && y + n <= R.h;
}

// One six-tap quarter-pel sample from the nine anchors, single final clip
// (reference src/motion.c:193-260).  dx, dy are wave-uniform.
__device__ __forceinline__ int q6_px(const PicDev &R, int x, int y, int dx, int dy)
{
    const uint8_t *yp = R.y + static_cast<size_t>(y) * R.pitch + x;
    const size_t ho = static_cast<size_t>(y) * R.hpitch + x;
    int top = 0, mid = 0, bot = 0;
    if (dy <= 1) {  // integer row y: i00, hh, i10
        switch (dx) {
        case 0: top = yp[0]; break;
        case 1: top = (yp[0] + R.hh[ho] + 1) >> 1; break;
        case 2: top = R.hh[ho]; break;
        default: top = (R.hh[ho] + yp[1] + 1) >> 1; break;
        }
    }
    if (dy >= 1) {  // half row: vv(x), cc, vv(x+1)
        switch (dx) {
        case 0: mid = R.vv[ho]; break;
        case 1: mid = (R.vv[ho] + R.cc[ho] + 1) >> 1; break;
        case 2: mid = R.cc[ho]; break;
        default: mid = (R.cc[ho] + R.vv[ho + 1] + 1) >> 1; break;
        }
    }
    if (dy == 3) {  // integer row y+1: i01, hh(y+1), i11
        const uint8_t *yq = yp + R.pitch;
        const size_t h1 = ho + R.hpitch;
        switch (dx) {
        case 0: bot = yq[0]; break;
        case 1: bot = (yq[0] + R.hh[h1] + 1) >> 1; break;
        case 2: bot = R.hh[h1]; break;
        default: bot = (R.hh[h1] + yq[1] + 1) >> 1; break;
        }
    }
    int v;
    switch (dy) {
    case 0: v = top; break;
    case 1: v = (top + mid + 1) >> 1; break;
    case 2: v = mid; break;
    default: v = (mid + bot + 1) >> 1; break;
    }
    return v < 0 ? 0 : (v > 255 ? 255 : v);
}

// Bilinear quarter-pel sample, rounded after each pass
// (reference src/motion.c:111-116, 330-336); never leaves [0,255].
__device__ __forceinline__ int bil_px(const PicDev &R, int x, int y, int dx, int dy)
{
    const uint8_t *p = R.y + static_cast<size_t>(y) * R.pitch + x;
    const int top = (p[0] * (4 - dx) + p[1] * dx + 2) >> 2;
    const int bot = (p[R.pitch] * (4 - dx) + p[R.pitch + 1] * dx + 2) >> 2;
    return (top * (4 - dy) + bot * dy + 2) >> 2;
}

// Quarter-pel candidate class (reference src/motion.c:455-464):
// 0 skipped, 1 six-tap, 2 bilinear.
__device__ __forceinline__ int qclass(const PicDev &R, int n, int qx, int qy)
{
    const int fx = qx >> 2, fy = qy >> 2;
    if (fx < 0 || fy < 0 || fx + n + 1 > R.w || fy + n + 1 > R.h) return 0;
    if (fx >= 2 && fy >= 2 && fx + n + 2 < R.w && fy + n + 2 < R.h) return 1;
    return 2;
}

// Part of one reference picture staged in LDS for a chunk of calls:
// frame region [x0,x1) x [y0,y1) at `lds`, `pitch` bytes per row (x0 is a
// multiple of 4; every row has 4 readable bytes past its end).
struct Win {
    int x0, y0, x1, y1, pitch;
    bool on;
};

template <int N>
struct Searcher {
    using Gm = Geo<N>;
    const PicDev &R;
    uint32_t o[Gm::U];    // current block, packed bytes per unit
    int urow[Gm::U], ucol[Gm::U];
    int roff[Gm::U];      // unit offset in the reference plane
    int grp;              // candidate group (N == 8)
    int o8;               // N == 8: this lane's current-block pixel
    gptr32 cplane;        // current picture
    size_t cblk;          // byte offset of the current block in it
    int cpitch;
    gptr32 rplane;        // reference picture
    Win W;                // staged reference window (W.on == false: none)
    Counts cnt;

    __device__ Searcher(const PicDev &ref, const PicDev &cur, int px, int py, const Win &win)
        : R(ref), cplane(as_global32(cur.y)), cblk(static_cast<size_t>(py) * cur.pitch + px),
          cpitch(cur.pitch), rplane(as_global32(ref.y)), W(win)
    {
        const int lane = threadIdx.x & 63;
        grp = Gm::G == 1 ? 0 : lane / Gm::UNITS;
#pragma unroll
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = Gm::G == 1 ? lane + 64 * u : lane % Gm::UNITS;
            urow[u] = unit / Gm::DW;
            ucol[u] = (unit % Gm::DW) * 4;
            o[u] = ld4g(cplane, cblk + static_cast<size_t>(urow[u]) * cpitch + ucol[u]);
            roff[u] = urow[u] * R.pitch + ucol[u];
        }
        o8 = 0;
        if (N == 8)
            o8 = cur.y[static_cast<size_t>(py + (lane >> 3)) * cur.pitch + px + (lane & 7)];
    }

    __device__ __forceinline__ bool in_win(int x, int y) const
    {
        return W.on && x >= W.x0 && y >= W.y0 && x + N <= W.x1 && y + N <= W.y1;
    }

    // Packed SAD of this lane's unit u for the block at (x,y): from the LDS
    // window when the block lies inside it, else from the reference plane.
    __device__ __forceinline__ uint32_t unit_sad(int x, int y, int u, bool lds, uint32_t acc) const
    {
        if (lds) {
            const int b = (y - W.y0) * W.pitch + (x - W.x0);
            return __builtin_amdgcn_sad_u8(ld4s(b + urow[u] * W.pitch + ucol[u]), o[u], acc);
        }
        const size_t b = static_cast<size_t>(y) * R.pitch + x;
        return __builtin_amdgcn_sad_u8(ld4g(rplane, b + roff[u]), o[u], acc);
    }

    // SAD of the block at full-pel (x,y), which must fit in the frame.
    __device__ __forceinline__ int32_t sad1(int x, int y) const
    {
        const bool lds = in_win(x, y);
        if (Gm::G == 1) {
            uint32_t acc = 0;
            if (lds) {
#pragma unroll
                for (int u = 0; u < Gm::U; ++u) acc = unit_sad(x, y, u, true, acc);
            } else {
#pragma unroll
                for (int u = 0; u < Gm::U; ++u) acc = unit_sad(x, y, u, false, acc);
            }
            return wave_sum(static_cast<int>(acc));
        } else {
            const int v = row_scan(static_cast<int>(unit_sad(x, y, 0, lds, 0u)));
            return __builtin_amdgcn_readlane(v, 15);  // group 0 priced it; all groups did
        }
    }

    // Prices candidates k >= first of a K-point set; others get kInf.
    template <int K>
    __device__ __forceinline__ void price_int(const int (&xs)[K], const int (&ys)[K], int first,
                                              int32_t (&s)[K]) const
    {
        if (Gm::G == 1) {
#pragma unroll
            for (int k = 0; k < K; ++k)
                s[k] = (k >= first && fits(R, N, xs[k], ys[k])) ? sad1(xs[k], ys[k]) : kInf;
        } else {
#pragma unroll
            for (int s0 = 0; s0 < K; s0 += 4) {
                if (s0 + 3 < first) {
#pragma unroll
                    for (int g = 0; g < 4; ++g)
                        if (s0 + g < K) s[s0 + g] = kInf;
                    continue;
                }
                int x = 0, y = 0;
                bool live = false;
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int k = s0 + g;
                    if (k < K && grp == g) {
                        live = k >= first && fits(R, N, xs[k], ys[k]);
                        x = xs[k];
                        y = ys[k];
                    }
                }
                if (!live) { x = 0; y = 0; }
                const int v = row_scan(static_cast<int>(unit_sad(x, y, 0, in_win(x, y), 0u)));
#pragma unroll
                for (int g = 0; g < 4; ++g) {
                    const int k = s0 + g;
                    if (k < K) {
                        const int r = __builtin_amdgcn_readlane(v, 16 * g + 15);
                        s[k] = (k >= first && fits(R, N, xs[k], ys[k])) ? r : kInf;
                    }
                }
            }
        }
    }

    // One greedy integer pattern stage (reference src/motion.c:399-448).
    // N >= 16: one candidate occupies the whole wave, so candidates are priced
    // in the reference's own order (no speculation, no wasted SADs).
    // N == 8: the four DPP rows price four candidates at once; the first
    // strict improvement in table order is taken and the rest re-priced.
    template <int P>
    __device__ __forceinline__ void stage(int scale, int passes, int &bx, int &by, int32_t &best)
    {
        constexpr int K = Pat<P>::K;
        if (Gm::G == 1) {
#pragma unroll 1
            for (int it = 0; it < passes; ++it) {
                bool moved = false;
#pragma unroll 1
                for (int k = 0; k < K; ++k) {
                    const int x = bx + Pat<P>::x(k) * scale, y = by + Pat<P>::y(k) * scale;
                    int32_t sv = kInf;
                    if (fits(R, N, x, y)) { sv = sad1(x, y); cnt.ie++; }
                    else cnt.oob++;
                    if (sv < best) { best = sv; bx = x; by = y; moved = true; }
                }
                if (!moved) break;
            }
            return;
        }
#pragma unroll 1
        for (int it = 0; it < passes; ++it) {
            bool moved = false;
            int first = 0;
            while (first < K) {
                int xs[K], ys[K];
                int32_t s[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    xs[k] = bx + Pat<P>::x(k) * scale;
                    ys[k] = by + Pat<P>::y(k) * scale;
                }
                price_int<K>(xs, ys, first, s);
                int hit = K, nx = bx, ny = by;
                int32_t nb = best;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (k >= first && hit == K) {
                        // the sequential loop reaches candidate k: count it
                        if (fits(R, N, xs[k], ys[k])) cnt.ie++; else cnt.oob++;
                        if (s[k] < best) { hit = k; nb = s[k]; nx = xs[k]; ny = ys[k]; }
                    }
                }
                if (hit == K) break;
                best = nb; bx = nx; by = ny; moved = true;
                first = hit + 1;
            }
            if (!moved) break;
        }
    }

    // SAD of the quarter-pel candidate (qx,qy) of class cls (1 or 2).
    __device__ __forceinline__ int32_t qsad(int qx, int qy, int cls) const
    {
        const int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
        if (N == 8) {
            const int lane = threadIdx.x & 63;
            const int x = fx + (lane & 7), y = fy + (lane >> 3);
            const int v = cls == 1 ? q6_px(R, x, y, dx, dy) : bil_px(R, x, y, dx, dy);
            const int d = o8 - v;
            return wave_sum(d < 0 ? -d : d);
        }
        // Units are walked with a rolled loop (registers stay bounded for
        // N = 64); the current block is re-read from its plane (cache hits)
        // so no register array is indexed dynamically.
        const int lane = threadIdx.x & 63;
        uint32_t acc = 0;
#pragma unroll 1
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = lane + 64 * u;
            const int ur = unit / Gm::DW, uc = (unit % Gm::DW) * 4;
            const int x = fx + uc, y = fy + ur;
            uint32_t pk = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int v = cls == 1 ? q6_px(R, x + j, y, dx, dy) : bil_px(R, x + j, y, dx, dy);
                pk |= static_cast<uint32_t>(v) << (8 * j);
            }
            acc = __builtin_amdgcn_sad_u8(pk, ld4g(cplane, cblk + static_cast<size_t>(ur) * cpitch + uc), acc);
        }
        return wave_sum(static_cast<int>(acc));
    }

    // Quarter-pel refinement, one greedy pass (reference src/motion.c:450-481).
    // Each candidate uses the whole wave, so they are priced in order.
    __device__ __forceinline__ void qpel(int &qx, int &qy, int32_t &best)
    {
        constexpr int K = Pat<PAT_QPEL>::K;
#pragma unroll 1
        for (int k = 0; k < K; ++k) {
            const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
            const int cls = qclass(R, N, tx, ty);
            if (!cls) continue;
            if (cls == 1) cnt.q6++; else cnt.qb++;
            const int32_t sv = qsad(tx, ty, cls);
            if (sv < best) { best = sv; qx = tx; qy = ty; }
        }
    }
};

template <int N>
__device__ void search(const PicDev &R, const PicDev &C, const tcme_call &c, tcme_result *res,
                       tcme_stats *st, const Win &win)
{
    Searcher<N> S(R, C, c.px, c.py, win);
    // Centre clamp with tc_clip's order: low bound first (motion.c:391-392).
    int bx = c.cx < 0 ? 0 : (c.cx > R.w - N ? R.w - N : c.cx);
    int by = c.cy < 0 ? 0 : (c.cy > R.h - N ? R.h - N : c.cy);
    int32_t best;
    if (fits(R, N, bx, by)) { best = S.sad1(bx, by); S.cnt.ie++; }
    else { best = kInf; S.cnt.oob++; }
    int qx, qy;
    if (R.w >= N && R.h >= N) {
        const int scale = c.sr / 4 > 1 ? c.sr / 4 : 1;
        S.template stage<PAT_LARGE>(scale, 2, bx, by, best);
        S.template stage<PAT_SMALL>(1, 3, bx, by, best);
        S.template stage<PAT_DIAMOND>(1, 4, bx, by, best);
        qx = bx * 4;
        qy = by * 4;
        S.qpel(qx, qy, best);
    } else {
        // No candidate can fit: one pass of each stage, all rejected; every
        // quarter-pel candidate fails the frame test.
        S.cnt.oob += 6 + 6 + 4;
        qx = bx * 4;
        qy = by * 4;
    }
    if ((threadIdx.x & 63) == 0) {
        res->mv_x = qx;
        res->mv_y = qy;
        res->sad = best;
        res->status = 0;
        if (st) {
            st->int_evals = S.cnt.ie;
            st->q6_evals = S.cnt.q6;
            st->qbil_evals = S.cnt.qb;
            st->int_oob = S.cnt.oob;
        }
    }
}

// Validates one descriptor; returns 0 and the two pictures, or a TCME_E_* code.
__device__ __forceinline__ int check_call(const PicDev *table, int npics, const tcme_call &c,
                                          PicDev &R, PicDev &C)
{
    if (c.ref_pic < 0 || c.ref_pic >= npics || c.cur_pic < 0 || c.cur_pic >= npics)
        return TCME_E_PIC;
    if (c.blk != 8 && c.blk != 16 && c.blk != 32 && c.blk != 64) return TCME_E_BLK;
    R = table[c.ref_pic];
    C = table[c.cur_pic];
    if (!R.y || !R.hh || !C.y) return TCME_E_PIC;  // empty slot, or no half-sample planes
    if (c.px < 0 || c.py < 0 || c.px + c.blk > C.w || c.py + c.blk > C.h) return TCME_E_ARG;
    return 0;
}

// Runs call i with the whole wave.
__device__ __forceinline__ void run_call(const PicDev *table, int npics, const tcme_call &c, int i,
                                         tcme_result *res, tcme_stats *stats, const Win &win)
{
    PicDev R{}, C{};
    const int status = check_call(table, npics, c, R, C);
    tcme_stats *st = stats ? stats + i : nullptr;
    if (status) {
        if ((threadIdx.x & 63) == 0) {
            res[i].mv_x = 0;
            res[i].mv_y = 0;
            res[i].sad = kInf;
            res[i].status = status;
            if (st) { st->int_evals = st->q6_evals = st->qbil_evals = st->int_oob = 0; }
        }
        return;
    }
    switch (c.blk) {
    case 8: search<8>(R, C, c, res + i, st, win); break;
    case 16: search<16>(R, C, c, res + i, st, win); break;
    case 32: search<32>(R, C, c, res + i, st, win); break;
    default: search<64>(R, C, c, res + i, st, win); break;
    }
}

// One wavefront per call, no staging (the single-call drop-in).
__global__ __launch_bounds__(64) void me_search_kernel(const PicDev *__restrict__ table, int npics,
                                                        const tcme_call *__restrict__ calls, int n,
                                                        tcme_result *__restrict__ res,
                                                        tcme_stats *__restrict__ stats)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const Win none{0, 0, 0, 0, 0, false};
    run_call(table, npics, calls[i], i, res, stats, none);
}

// Block index -> chunk index so that each of the 8 XCDs (which receive
// workgroups round-robin) works on one contiguous run of chunks: neighbouring
// chunks read neighbouring reference rows, which then stay in that XCD's L2.
__device__ __forceinline__ int xcd_chunk(int b, int nb)
{
    constexpr int kXcd = 8;
    const int x = b % kXcd, slot = b / kXcd;
    const int q = nb / kXcd, r = nb % kXcd;
    return x * q + (x < r ? x : r) + slot;
}

// Relative cost of one call, for splitting a chunk between wavefronts.
__device__ __forceinline__ int call_cost(int blk)
{
    return blk * blk + 256;
}

// Chunked search: a workgroup of WAVES wavefronts takes `chunk` consecutive
// calls and stages the union of their reference windows (clamped centre +-
// apron, one reference picture) in LDS with coalesced dword loads.  Each
// wavefront then takes a contiguous, cost-balanced share of the chunk (calls
// ordered largest first split evenly) and prices candidates from LDS; a
// candidate outside the window reads the reference plane instead, so the
// apron only affects speed, never results.  All control flow below the
// barriers is wave-uniform: the wavefront index comes from readfirstlane and
// the share from scalar loads of the call list.
template <int WAVES>
__global__ __launch_bounds__(64 * WAVES) void me_tile_kernel(
    const PicDev *__restrict__ table, int npics, const tcme_call *__restrict__ calls, int n,
    int chunk, int apron, int lds_cap, tcme_result *__restrict__ res,
    tcme_stats *__restrict__ stats)
{
    __shared__ int box[5];
    const int tid = threadIdx.x;
    const int b = xcd_chunk(blockIdx.x, gridDim.x);
    const int c0 = b * chunk;
    const int c1 = min(n, c0 + chunk);
    if (c0 >= n) return;  // uniform for the whole workgroup
    if (tid == 0) {
        box[0] = 0x7FFFFFFF; box[1] = 0x7FFFFFFF;
        box[2] = -0x7FFFFFFF; box[3] = -0x7FFFFFFF;
        box[4] = calls[c0].ref_pic;
    }
    __syncthreads();
    const int ref0 = box[4];
    for (int i = c0 + tid; i < c1; i += 64 * WAVES) {
        const tcme_call c = calls[i];
        PicDev R{}, C{};
        if (c.ref_pic != ref0 || check_call(table, npics, c, R, C)) continue;
        const int n_ = c.blk;
        const int cx = c.cx < 0 ? 0 : (c.cx > R.w - n_ ? R.w - n_ : c.cx);
        const int cy = c.cy < 0 ? 0 : (c.cy > R.h - n_ ? R.h - n_ : c.cy);
        atomicMin(&box[0], cx - apron);
        atomicMin(&box[1], cy - apron);
        atomicMax(&box[2], cx + n_ + apron);
        atomicMax(&box[3], cy + n_ + apron);
    }
    __syncthreads();
    Win win{0, 0, 0, 0, 0, false};
    if (box[0] != 0x7FFFFFFF) {
        const PicDev R = table[ref0];
        const int x0 = (box[0] < 0 ? 0 : box[0]) & ~3;
        const int y0 = box[1] < 0 ? 0 : box[1];
        const int x1 = box[2] > R.w ? R.w : box[2];
        const int y1 = box[3] > R.h ? R.h : box[3];
        const int pitch = ((x1 - x0 + 3) & ~3) + 4;
        const int rows = y1 - y0;
        if (x1 > x0 && rows > 0 && rows * pitch + 8 <= lds_cap) {
            const int words = pitch / 4;
            gptr32 src = as_global32(R.y);
            for (int i = tid; i < rows * words; i += 64 * WAVES) {
                const int r = i / words, cw = i - r * words;
                tcme_win[i] = src[(static_cast<size_t>(y0 + r) * R.pitch + x0) / 4 + cw];
            }
            if (tid == 0) tcme_win[rows * words] = 0;  // the pair read past the last row
            win = Win{x0, y0, x1, y1, pitch, true};
        }
    }
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    int total = 0;
    for (int j = c0; j < c1; ++j) total += call_cost(calls[j].blk);
    const int lo = static_cast<int>(static_cast<long long>(total) * wave / WAVES);
    const int hi = static_cast<int>(static_cast<long long>(total) * (wave + 1) / WAVES);
    int acc = 0;
    for (int j = c0; j < c1; ++j) {
        const tcme_call c = calls[j];
        const int cj = call_cost(c.blk);
        const int mid = acc + cj / 2;
        acc += cj;
        if (mid < lo || mid >= hi) continue;
        Win w = win;
        w.on = win.on && c.ref_pic == ref0;
        run_call(table, npics, c, j, res, stats, w);
    }
}

// Raw six-tap half-sample planes (reference src/motion.c:70-107), one thread
// per sample; positions whose taps would leave the frame are written as 0.
__global__ __launch_bounds__(256) void halfpel_kernel(const uint8_t *__restrict__ y, int pitch,
                                                      int w, int h, int16_t *__restrict__ hh,
                                                      int16_t *__restrict__ vv,
                                                      int16_t *__restrict__ cc, int hpitch)
{
    const int x = blockIdx.x * 64 + threadIdx.x;
    const int r = blockIdx.y * 4 + threadIdx.y;
    if (x >= w || r >= h) return;
    const bool okx = x >= 2 && x + 3 < w;
    const bool oky = r >= 2 && r + 3 < h;
    int hv = 0, vval = 0, cval = 0;
    const uint8_t *row = y + static_cast<size_t>(r) * pitch;
    if (okx) hv = tap6(row[x - 2], row[x - 1], row[x], row[x + 1], row[x + 2], row[x + 3]);
    if (oky) {
        const uint8_t *col = row + x;
        vval = tap6(col[-2 * pitch], col[-pitch], col[0], col[pitch], col[2 * pitch], col[3 * pitch]);
    }
    if (okx && oky) {
        int t[6];
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const uint8_t *rr = y + static_cast<size_t>(r - 2 + k) * pitch + x;
            t[k] = tap6(rr[-2], rr[-1], rr[0], rr[1], rr[2], rr[3]);
        }
        cval = tap6(t[0], t[1], t[2], t[3], t[4], t[5]);
    }
    const size_t o = static_cast<size_t>(r) * hpitch + x;
    hh[o] = static_cast<int16_t>(hv);
    vv[o] = static_cast<int16_t>(vval);
    cc[o] = static_cast<int16_t>(cval);
}

}  // namespace

hipError_t launch_halfpel(const PicDev &p, hipStream_t s)
{
    dim3 grid((p.w + 63) / 64, (p.h + 3) / 4), block(64, 4);
    hipLaunchKernelGGL(halfpel_kernel, grid, block, 0, s, p.y, p.pitch, p.w, p.h,
                       const_cast<int16_t *>(p.hh), const_cast<int16_t *>(p.vv),
                       const_cast<int16_t *>(p.cc), p.hpitch);
    return hipGetLastError();
}

hipError_t launch_search(const PicDev *table, int npics, const tcme_call *calls, int n,
                         tcme_result *res, tcme_stats *stats, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(me_search_kernel, dim3(n), dim3(64), 0, s, table, npics, calls, n, res,
                       stats);
    return hipGetLastError();
}

hipError_t launch_search_tiled(const PicDev *table, int npics, const tcme_call *calls, int n,
                               const Tiling &t, tcme_result *res, tcme_stats *stats,
                               hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    if (t.chunk == 1 && t.lds_bytes == 0)  // the plain form: one wavefront per call
        return launch_search(table, npics, calls, n, res, stats, s);
    const int nb = (n + t.chunk - 1) / t.chunk;
    switch (t.waves) {
    case 1:
        hipLaunchKernelGGL(me_tile_kernel<1>, dim3(nb), dim3(64), t.lds_bytes, s, table, npics,
                           calls, n, t.chunk, t.apron, t.lds_bytes, res, stats);
        break;
    case 2:
        hipLaunchKernelGGL(me_tile_kernel<2>, dim3(nb), dim3(128), t.lds_bytes, s, table, npics,
                           calls, n, t.chunk, t.apron, t.lds_bytes, res, stats);
        break;
    case 8:
        hipLaunchKernelGGL(me_tile_kernel<8>, dim3(nb), dim3(512), t.lds_bytes, s, table, npics,
                           calls, n, t.chunk, t.apron, t.lds_bytes, res, stats);
        break;
    default:
        hipLaunchKernelGGL(me_tile_kernel<4>, dim3(nb), dim3(256), t.lds_bytes, s, table, npics,
                           calls, n, t.chunk, t.apron, t.lds_bytes, res, stats);
        break;
    }
    return hipGetLastError();
}

}  // namespace tcme
