// gfx950 kernels for TCodec's motion-estimation path.
//
//   anchor_kernel     per reference picture: every sample with its raw six-tap
//                     half samples, packed 8 bytes per sample (reference
//                     src/motion.c:70-107)
//   me_search_kernel  the hierarchical hexagon search + quarter-pel refinement,
//                     one wavefront per call (reference src/motion.c:384-485)
//   me_tile_kernel    the same search for a chunk of calls per workgroup, with
//                     the union of their reference windows staged in LDS
//
// The walk is the reference's sequential greedy walk.  Its decisions are
// wave-uniform (scalar unit); lanes price candidates.  Candidates may be
// priced speculatively in batches around the current best: the FIRST one in
// table order that is strictly better is taken and those after it are priced
// again around the new best.  That is exactly what the sequential loop
// computes, because the candidates before the first improvement saw the same
// centre, so results are bit-identical whatever the batch size.
//
// Integer block costs: v_sad_u8 on packed bytes (4 |a-b| per lane-op) and a
// DPP row scan.  Quarter-pel: every candidate of one call lies within one
// sample of the integer best (fx in {X-1, X}, fy in {Y-1, Y}), so a lane loads
// its 3-row anchor neighbourhood once and prices candidates from registers.
#include "tcme_internal.h"

namespace tcme {
namespace {

constexpr int32_t kInf = 0x7FFFFFFF;

__device__ __forceinline__ int tap6(int a, int b, int c, int d, int e, int f)
{
    return (a + f - 5 * (b + e) + 20 * (c + d) + 16) >> 5;
}

// Global-memory views (address space 1), so plane reads compile to global
// loads rather than flat ones.
typedef const __attribute__((address_space(1))) uint32_t *gptr32;
typedef const __attribute__((address_space(1))) uint64_t *gptr64;

__device__ __forceinline__ gptr32 as_global32(const void *p) { return (gptr32)(p); }
__device__ __forceinline__ gptr64 as_global64(const void *p) { return (gptr64)(p); }

// One anchor entry: bits 0-15 sample, 16-31 hh, 32-47 vv, 48-63 cc.
typedef uint64_t anc_t;

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
// N >= 16: every lane owns U units of one candidate (G == 1); integer
//          candidates are priced B at a time.
// N == 8 : 16 units, so the wave's four DPP rows price four candidates (G == 4).
template <int N> struct Geo {
    static constexpr int DW = N / 4;
    static constexpr int UNITS = N * DW;
    static constexpr int U = UNITS >= 64 ? UNITS / 64 : 1;
    static constexpr int G = UNITS >= 64 ? 1 : 64 / UNITS;
    static constexpr int B = N >= 64 ? 2 : 6;
};

struct Counts { int ie = 0, oob = 0, q6 = 0, qb = 0; };

__device__ __forceinline__ bool fits(const PicDev &R, int n, int x, int y)
{
    return x >= 0 && y >= 0 && x + n <= R.w && y + n <= R.h;
}

// ---- anchors ---------------------------------------------------------------
// Entry (x,y): sample i(x,y), then hh/vv/cc, the raw six-tap half samples
// right of, below and diagonal to the sample (all fit int16).
__device__ __forceinline__ int a_i(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e) & 0xFFFFu); }
__device__ __forceinline__ int a_h(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e)) >> 16; }
__device__ __forceinline__ int a_v(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e >> 32) << 16) >> 16; }
__device__ __forceinline__ int a_c(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e >> 32)) >> 16; }

// Quarter-sample value from the four anchor entries around it, raw
// intermediates and a single final clip (reference src/motion.c:193-260).
// Each anchor row is a lattice L0 (sample), L1 (half), L2 (next sample); a
// phase p picks L0, avg(L0,L1), L1, avg(L1,L2); with a == b for even phases,
// (a + b + 1) >> 1 == a, so one expression covers all four.
__device__ __forceinline__ int lattice(int p, int l0, int l1, int l2)
{
    const int a = p >= 2 ? l1 : l0;
    const int b = p == 1 ? l1 : (p == 3 ? l2 : a);
    return (a + b + 1) >> 1;
}

__device__ __forceinline__ int qpix(anc_t p00, anc_t p10, anc_t p01, anc_t p11, int dx, int dy)
{
    const int top = lattice(dx, a_i(p00), a_h(p00), a_i(p10));
    const int mid = lattice(dx, a_v(p00), a_c(p00), a_v(p10));
    const int bot = lattice(dx, a_i(p01), a_h(p01), a_i(p11));
    const int v = lattice(dy, top, mid, bot);
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
// frame region [x0,x1) x [y0,y1), `pitch` bytes per row (x0 is a multiple of
// 4; every row has 4 readable bytes past its end).
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
    int grp;              // candidate group (N == 8)
    int o8;               // N == 8: this lane's current-block pixel
    gptr32 cplane;        // current picture
    size_t cblk;          // byte offset of the current block in it
    int cpitch;
    gptr32 rplane;        // reference picture
    gptr64 aplane;        // reference anchors
    Win W;                // staged reference window (W.on == false: none)
    Counts cnt;

    __device__ Searcher(const PicDev &ref, const PicDev &cur, int px, int py, const Win &win)
        : R(ref), cplane(as_global32(cur.y)), cblk(static_cast<size_t>(py) * cur.pitch + px),
          cpitch(cur.pitch), rplane(as_global32(ref.y)), aplane(as_global64(ref.an)), W(win)
    {
        const int lane = threadIdx.x & 63;
        grp = Gm::G == 1 ? 0 : lane / Gm::UNITS;
#pragma unroll
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = Gm::G == 1 ? lane + 64 * u : lane % Gm::UNITS;
            urow[u] = unit / Gm::DW;
            ucol[u] = (unit % Gm::DW) * 4;
            o[u] = ld4g(cplane, cblk + static_cast<size_t>(urow[u]) * cpitch + ucol[u]);
        }
        o8 = 0;
        if (N == 8) {
            const uint32_t w = ld4g(cplane, cblk + static_cast<size_t>(lane >> 3) * cpitch + (lane & 4));
            o8 = (w >> (8 * (lane & 3))) & 0xFF;
        }
    }

    __device__ __forceinline__ bool in_win(int x, int y) const
    {
        return W.on && x >= W.x0 && y >= W.y0 && x + N <= W.x1 && y + N <= W.y1;
    }

    // Reference dword of unit u for the block at (x,y): from the LDS window
    // when the block lies inside it, else from the reference plane.
    __device__ __forceinline__ uint32_t ref4(int x, int y, int u, bool lds) const
    {
        if (lds) return ld4s((y - W.y0 + urow[u]) * W.pitch + (x - W.x0) + ucol[u]);
        return ld4g(rplane, static_cast<size_t>(y + urow[u]) * R.pitch + x + ucol[u]);
    }

    // SAD of the block at full-pel (x,y), which must fit in the frame.
    __device__ __forceinline__ int32_t sad1(int x, int y) const
    {
        const bool lds = in_win(x, y);
        if (Gm::G == 1) {
            uint32_t acc = 0;
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) acc = __builtin_amdgcn_sad_u8(ref4(x, y, u, lds), o[u], acc);
            return wave_sum(static_cast<int>(acc));
        }
        const int v = row_scan(static_cast<int>(__builtin_amdgcn_sad_u8(ref4(x, y, 0, lds), o[0], 0u)));
        return __builtin_amdgcn_readlane(v, 15);  // group 0 priced it; all groups did
    }

    // N >= 16: prices candidates first .. first+B-1 of pattern P (those < K)
    // around (bx,by); all their loads are issued before the first SAD so the
    // batch costs one memory latency.  Off-frame ones get kInf.
    template <int P>
    __device__ __forceinline__ void price_batch(int bx, int by, int scale, int first,
                                                int32_t (&s)[Gm::B]) const
    {
        constexpr int K = Pat<P>::K;
        uint32_t v[Gm::B][Gm::U];
        bool ok[Gm::B];
#pragma unroll
        for (int b = 0; b < Gm::B; ++b) {
            const int k = first + b;
            const int x = bx + Pat<P>::x(k) * scale, y = by + Pat<P>::y(k) * scale;
            ok[b] = k < K && fits(R, N, x, y);
            // a rejected candidate reads the (in-frame) centre instead
            const int lx = ok[b] ? x : bx, ly = ok[b] ? y : by;
            const bool lds = in_win(lx, ly);
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) v[b][u] = ref4(lx, ly, u, lds);
        }
#pragma unroll
        for (int b = 0; b < Gm::B; ++b) {
            uint32_t acc = 0;
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) acc = __builtin_amdgcn_sad_u8(v[b][u], o[u], acc);
            const int t = wave_sum(static_cast<int>(acc));
            s[b] = ok[b] ? t : kInf;
        }
    }

    // N == 8: prices candidates k >= first of a K-point set, four per DPP
    // row group; others get kInf.
    template <int K>
    __device__ __forceinline__ void price_grouped(const int (&xs)[K], const int (&ys)[K], int first,
                                                  int32_t (&s)[K]) const
    {
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
            const int v = row_scan(static_cast<int>(__builtin_amdgcn_sad_u8(ref4(x, y, 0, in_win(x, y)), o[0], 0u)));
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

    // One greedy integer pattern stage (reference src/motion.c:399-448).
    template <int P>
    __device__ __forceinline__ void stage(int scale, int passes, int &bx, int &by, int32_t &best)
    {
        constexpr int K = Pat<P>::K;
        if (Gm::G == 1) {
#pragma unroll 1
            for (int it = 0; it < passes; ++it) {
                bool moved = false;
                int first = 0;
#pragma unroll 1
                while (first < K) {
                    int32_t s[Gm::B];
                    price_batch<P>(bx, by, scale, first, s);
                    int hit = -1;
#pragma unroll
                    for (int b = 0; b < Gm::B; ++b) {
                        const int k = first + b;
                        if (k < K && hit < 0) {
                            // the sequential loop reaches candidate k: count it
                            const int x = bx + Pat<P>::x(k) * scale, y = by + Pat<P>::y(k) * scale;
                            if (fits(R, N, x, y)) cnt.ie++; else cnt.oob++;
                            if (s[b] < best) hit = b;
                        }
                    }
                    if (hit < 0) { first += Gm::B; continue; }
                    int32_t nb = kInf;
#pragma unroll
                    for (int b = 0; b < Gm::B; ++b)
                        if (b == hit) nb = s[b];
                    const int k = first + hit;
                    bx += Pat<P>::x(k) * scale;
                    by += Pat<P>::y(k) * scale;
                    best = nb;
                    moved = true;
                    first = k + 1;
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
                price_grouped<K>(xs, ys, first, s);
                int hit = K, nx = bx, ny = by;
                int32_t nb = best;
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (k >= first && hit == K) {
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

    // ---- quarter-pel --------------------------------------------------------

    // SAD of a quarter-pel candidate read directly from the planes (edge path).
    __device__ __forceinline__ int32_t qsad_direct(int qx, int qy, int cls) const
    {
        const int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
        const int lane = threadIdx.x & 63;
        if (N == 8) {
            const int x = fx + (lane & 7), y = fy + (lane >> 3);
            int v;
            if (cls == 1) {
                const size_t e = static_cast<size_t>(y) * R.apitch + x;
                v = qpix(aplane[e], aplane[e + 1], aplane[e + R.apitch], aplane[e + R.apitch + 1], dx, dy);
            } else {
                v = bil_px(R, x, y, dx, dy);
            }
            const int d = o8 - v;
            return wave_sum(d < 0 ? -d : d);
        }
        uint32_t acc = 0;
#pragma unroll 1
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = lane + 64 * u;
            const int ur = unit / Gm::DW, uc = (unit % Gm::DW) * 4;
            const int x = fx + uc, y = fy + ur;
            uint32_t pk = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int v;
                if (cls == 1) {
                    const size_t e = static_cast<size_t>(y) * R.apitch + x + j;
                    v = qpix(aplane[e], aplane[e + 1], aplane[e + R.apitch], aplane[e + R.apitch + 1], dx, dy);
                } else {
                    v = bil_px(R, x + j, y, dx, dy);
                }
                pk |= static_cast<uint32_t>(v) << (8 * j);
            }
            acc = __builtin_amdgcn_sad_u8(pk, ld4g(cplane, cblk + static_cast<size_t>(ur) * cpitch + uc), acc);
        }
        return wave_sum(static_cast<int>(acc));
    }

    // Packed SAD of one unit (4 pixels) for a candidate whose anchor rows are
    // ra (sample row) and rb (row below), starting at column ox.
    template <int C>
    __device__ __forceinline__ uint32_t unit_q(const anc_t (&ra)[C], const anc_t (&rb)[C], int ox,
                                               int dx, int dy, uint32_t orig, uint32_t acc) const
    {
        uint32_t pk = 0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const anc_t p00 = ox ? ra[j + 1] : ra[j], p10 = ox ? ra[j + 2] : ra[j + 1];
            const anc_t p01 = ox ? rb[j + 1] : rb[j], p11 = ox ? rb[j + 2] : rb[j + 1];
            pk |= static_cast<uint32_t>(qpix(p00, p10, p01, p11, dx, dy)) << (8 * j);
        }
        return __builtin_amdgcn_sad_u8(pk, orig, acc);
    }

    // Quarter-pel refinement, one greedy pass (reference src/motion.c:450-481).
    // (bx,by) is the integer best.  When every sample position the pass can
    // reach (fx in {bx-1, bx}, fy in {by-1, by}) is six-tap safe, the anchor
    // neighbourhood is loaded once: N <= 16 keep it in registers and price
    // candidates in the reference's order; N >= 32 price all remaining
    // candidates per unit in rounds (first strict improvement wins, the rest
    // are re-priced).  Otherwise candidates are priced one by one from the
    // planes with their own class.
    __device__ __forceinline__ void qpel(int bx, int by, int &qx, int &qy, int32_t &best)
    {
        constexpr int K = Pat<PAT_QPEL>::K;
        qx = bx * 4;
        qy = by * 4;
        const bool fast = bx - 1 >= 2 && by - 1 >= 2 && bx + N + 2 < R.w && by + N + 2 < R.h;
        if (!fast) {
#pragma unroll 1
            for (int k = 0; k < K; ++k) {
                const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
                const int cls = qclass(R, N, tx, ty);
                if (!cls) continue;
                if (cls == 1) cnt.q6++; else cnt.qb++;
                const int32_t sv = qsad_direct(tx, ty, cls);
                if (sv < best) { best = sv; qx = tx; qy = ty; }
            }
            return;
        }
        const int lane = threadIdx.x & 63;
        const int x0 = bx - 1, y0 = by - 1;  // window origin (sample coordinates)
        if (N == 8) {
            // lane = pixel; a 3x3 anchor window in registers
            const int c = lane & 7, r = lane >> 3;
            anc_t e[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j)
                    e[i][j] = aplane[static_cast<size_t>(y0 + r + i) * R.apitch + x0 + c + j];
#pragma unroll 1
            for (int k = 0; k < K; ++k) {
                const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
                const int ox = (tx >> 2) - x0, oy = (ty >> 2) - y0;
                const int dx = tx & 3, dy = ty & 3;
                cnt.q6++;
                const anc_t p00 = oy ? (ox ? e[1][1] : e[1][0]) : (ox ? e[0][1] : e[0][0]);
                const anc_t p10 = oy ? (ox ? e[1][2] : e[1][1]) : (ox ? e[0][2] : e[0][1]);
                const anc_t p01 = oy ? (ox ? e[2][1] : e[2][0]) : (ox ? e[1][1] : e[1][0]);
                const anc_t p11 = oy ? (ox ? e[2][2] : e[2][1]) : (ox ? e[1][2] : e[1][1]);
                const int d = o8 - qpix(p00, p10, p01, p11, dx, dy);
                const int32_t sv = wave_sum(d < 0 ? -d : d);
                if (sv < best) { best = sv; qx = tx; qy = ty; }
            }
            return;
        }
        if (Gm::U == 1) {
            // N == 16: one unit per lane; a 3x6 anchor window in registers
            anc_t e[3][6];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 6; ++j)
                    e[i][j] = aplane[static_cast<size_t>(y0 + urow[0] + i) * R.apitch + x0 + ucol[0] + j];
#pragma unroll 1
            for (int k = 0; k < K; ++k) {
                const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
                const int ox = (tx >> 2) - x0, oy = (ty >> 2) - y0;
                cnt.q6++;
                uint32_t acc;
                if (oy) acc = unit_q<6>(e[1], e[2], ox, tx & 3, ty & 3, o[0], 0u);
                else acc = unit_q<6>(e[0], e[1], ox, tx & 3, ty & 3, o[0], 0u);
                const int32_t sv = wave_sum(static_cast<int>(acc));
                if (sv < best) { best = sv; qx = tx; qy = ty; }
            }
            return;
        }
        // N >= 32: rounds over the units, all remaining candidates per unit
        int first = 0;
#pragma unroll 1
        while (first < K) {
            uint32_t acc[K];
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] = 0;
#pragma unroll 1
            for (int u = 0; u < Gm::U; ++u) {
                const int unit = lane + 64 * u;
                const int ur = unit / Gm::DW, uc = (unit % Gm::DW) * 4;
                anc_t e[3][6];
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 6; ++j)
                        e[i][j] = aplane[static_cast<size_t>(y0 + ur + i) * R.apitch + x0 + uc + j];
                const uint32_t orig = ld4g(cplane, cblk + static_cast<size_t>(ur) * cpitch + uc);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if (k < first) continue;
                    const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
                    const int ox = (tx >> 2) - x0, oy = (ty >> 2) - y0;
                    if (oy) acc[k] = unit_q<6>(e[1], e[2], ox, tx & 3, ty & 3, orig, acc[k]);
                    else acc[k] = unit_q<6>(e[0], e[1], ox, tx & 3, ty & 3, orig, acc[k]);
                }
            }
            int hit = K, nx = qx, ny = qy;
            int32_t nb = best;
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if (k >= first && hit == K) {
                    cnt.q6++;
                    const int32_t sv = wave_sum(static_cast<int>(acc[k]));
                    if (sv < best) {
                        hit = k;
                        nb = sv;
                        nx = qx + Pat<PAT_QPEL>::x(k);
                        ny = qy + Pat<PAT_QPEL>::y(k);
                    }
                }
            }
            if (hit == K) break;
            best = nb; qx = nx; qy = ny;
            first = hit + 1;
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
        S.qpel(bx, by, qx, qy, best);
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
    if (!R.y || !R.an || !C.y) return TCME_E_PIC;  // empty slot, or a reference without anchors
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
#ifndef TCME_ONLY_N
    case 8: search<8>(R, C, c, res + i, st, win); break;
    case 16: search<16>(R, C, c, res + i, st, win); break;
    case 32: search<32>(R, C, c, res + i, st, win); break;
    default: search<64>(R, C, c, res + i, st, win); break;
#else
    default: search<TCME_ONLY_N>(R, C, c, res + i, st, win); break;
#endif
    }
}

// One wavefront per call, block size N only: the plain form launches one
// instance per size (their register budgets differ 4x: an 8x8 search needs
// about a quarter of a 64x64 one, so small blocks run at full occupancy).
// A wave whose call has another size exits at once; the 64 instance also
// answers descriptors with an unsupported size.
template <int N>
__global__ __launch_bounds__(64) void me_search_kernel(const PicDev *__restrict__ table, int npics,
                                                        const tcme_call *__restrict__ calls, int n,
                                                        tcme_result *__restrict__ res,
                                                        tcme_stats *__restrict__ stats)
{
    const int i = blockIdx.x;
    if (i >= n) return;
    const tcme_call c = calls[i];
    const bool mine = c.blk == N || (N == 64 && c.blk != 8 && c.blk != 16 && c.blk != 32);
    if (!mine) return;
    PicDev R{}, C{};
    const int status = check_call(table, npics, c, R, C);
    tcme_stats *st = stats ? stats + i : nullptr;
    if (status) {
        if (threadIdx.x == 0) {
            res[i].mv_x = 0;
            res[i].mv_y = 0;
            res[i].sad = kInf;
            res[i].status = status;
            if (st) { st->int_evals = st->q6_evals = st->qbil_evals = st->int_oob = 0; }
        }
        return;
    }
    const Win none{0, 0, 0, 0, 0, false};
    search<N>(R, C, c, res + i, st, none);
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
// wavefront then takes a contiguous, cost-balanced share of the chunk and
// prices integer candidates from LDS; a candidate outside the window reads
// the reference plane instead, so the apron only affects speed, never
// results.  All control flow below the barriers is wave-uniform: the
// wavefront index comes from readfirstlane and the share from scalar loads of
// the call list.
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

// Anchor plane of a picture (reference src/motion.c:70-107 for the taps):
// one thread per sample writes {sample, hh, vv, cc}; half samples whose taps
// would leave the frame are written as 0 and never read by the search.
__global__ __launch_bounds__(256) void anchor_kernel(const uint8_t *__restrict__ y, int pitch,
                                                     int w, int h, uint64_t *__restrict__ an,
                                                     int apitch)
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
    const uint32_t lo = static_cast<uint32_t>(row[x]) | (static_cast<uint32_t>(hv) << 16);
    const uint32_t hi = (static_cast<uint32_t>(vval) & 0xFFFFu) | (static_cast<uint32_t>(cval) << 16);
    an[static_cast<size_t>(r) * apitch + x] = static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
}

}  // namespace

hipError_t launch_anchors(const PicDev &p, hipStream_t s)
{
    dim3 grid((p.w + 63) / 64, (p.h + 3) / 4), block(64, 4);
    hipLaunchKernelGGL(anchor_kernel, grid, block, 0, s, p.y, p.pitch, p.w, p.h,
                       const_cast<uint64_t *>(p.an), p.apitch);
    return hipGetLastError();
}

hipError_t launch_search_n(int blk, const PicDev *table, int npics, const tcme_call *calls, int n,
                           tcme_result *res, tcme_stats *stats, hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    switch (blk) {
    case 8:
        hipLaunchKernelGGL(me_search_kernel<8>, dim3(n), dim3(64), 0, s, table, npics, calls, n, res, stats);
        break;
    case 16:
        hipLaunchKernelGGL(me_search_kernel<16>, dim3(n), dim3(64), 0, s, table, npics, calls, n, res, stats);
        break;
    case 32:
        hipLaunchKernelGGL(me_search_kernel<32>, dim3(n), dim3(64), 0, s, table, npics, calls, n, res, stats);
        break;
    default:
        hipLaunchKernelGGL(me_search_kernel<64>, dim3(n), dim3(64), 0, s, table, npics, calls, n, res, stats);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_search_tiled(const PicDev *table, int npics, const tcme_call *calls, int n,
                               const Tiling &t, tcme_result *res, tcme_stats *stats,
                               hipStream_t s)
{
    if (n <= 0) return hipSuccess;
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
