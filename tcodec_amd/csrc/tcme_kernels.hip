// gfx950 kernels for TCodec's motion-estimation path.
//
//   anchor_kernel     per reference picture: every sample with its raw six-tap
//                     half samples, packed 8 bytes per sample (reference
//                     src/motion.c:70-107)
//   me_tile_kernel    the hierarchical hexagon search + quarter-pel refinement
//                     (reference src/motion.c:384-485) for many calls per
//                     workgroup: runs of calls that share a reference picture
//                     and a 64x64 centre tile search from a luma window and a
//                     half-sample lattice staged in LDS once, each wavefront
//                     walking several calls side by side (L lanes per call).
//                     The <8, true> instance takes every block size and is
//                     the default: one launch per frame.
//   me_search_kernel  the same search, one workgroup per call, WPC wavefronts
//                     splitting its block, one instance per block size: the
//                     drop-in's single calls and the other launch shapes.
//
// The walk is the reference's sequential greedy walk.  In the tile kernel
// each group of L lanes runs it for its own call on vector registers.  In the
// per-call kernel its decisions are wave-uniform (scalar unit) and lanes price
// candidates, possibly in batches around the current best: the FIRST one in
// table order that is strictly better is taken and those after it are priced
// again around the new best.  That is exactly what the sequential loop
// computes, because the candidates before the first improvement saw the same
// centre, so results are bit-identical whatever the batch size.
//
// Integer block costs: v_sad_u8 on packed bytes (4 |a-b| per lane-op) and DPP
// sums (an LDS exchange when several wavefronts share a block).  Quarter-pel:
// every candidate of one call lies within one sample of the integer best, so
// the six-tap values come from the run's lattice or from an (N+2) x (N+2)
// anchor window staged once in LDS.
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
typedef const __attribute__((address_space(1))) uint8_t *gptr8;
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef const __attribute__((address_space(1))) u32x4 *gptr128;
typedef const __attribute__((address_space(1))) uint32_t *gptr32;
typedef const __attribute__((address_space(1))) uint64_t *gptr64;

__device__ __forceinline__ gptr8 as_global8(const void *p) { return (gptr8)(p); }
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

struct Counts { int ie = 0, oob = 0, q6 = 0, qb = 0; };

__device__ __forceinline__ bool fits(const PicDev &R, int n, int x, int y)
{
    return block_fits(x, y, n, R.w, R.h);
}

// ---- anchors ---------------------------------------------------------------
// Entry (x,y): sample i(x,y), then hh/vv/cc, the raw six-tap half samples
// right of, below and diagonal to the sample (all fit int16).
__device__ __forceinline__ int a_i(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e) & 0xFFFFu); }
__device__ __forceinline__ int a_h(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e)) >> 16; }
__device__ __forceinline__ int a_v(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e >> 32) << 16) >> 16; }
__device__ __forceinline__ int a_c(anc_t e) { return static_cast<int>(static_cast<uint32_t>(e >> 32)) >> 16; }

// Bilinear quarter-pel sample, rounded after each pass
// (reference src/motion.c:111-116, 330-336); never leaves [0,255].
__device__ __forceinline__ int bil_px(const PicDev &R, int x, int y, int dx, int dy)
{
    const gptr8 p = as_global8(R.y) + static_cast<size_t>(y) * R.pitch + x;
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


// Quarter-sample value from the anchors around it for phase (DX, DY), raw
// intermediates and a single final clip (reference src/motion.c:193-260).
// Each anchor row is a lattice L0 (sample), L1 (half), L2 (next sample);
// phase p picks L0, avg(L0,L1), L1 or avg(L1,L2), and the rows are blended
// the same way by DY.  Only the anchors the phase uses are read.
template <int P>
__device__ __forceinline__ int lat(int l0, int l1, int l2)
{
    if (P == 0) return l0;
    if (P == 1) return (l0 + l1 + 1) >> 1;
    if (P == 2) return l1;
    return (l1 + l2 + 1) >> 1;
}

// (E: an LDS anchor pointer or a global one, gptr64.)
template <int DX, int DY, typename E>
__device__ __forceinline__ int qpix(E e, int stride)
{
    // e points at the sample's entry; e[1] is right of it, e[stride] below
    const anc_t p00 = e[0];
    const anc_t p10 = (DX == 3) ? e[1] : p00;
    int v;
    if (DY == 0) {
        v = lat<DX>(a_i(p00), a_h(p00), a_i(p10));
    } else if (DY == 2) {
        v = lat<DX>(a_v(p00), a_c(p00), a_v(p10));
    } else if (DY == 1) {
        v = (lat<DX>(a_i(p00), a_h(p00), a_i(p10)) + lat<DX>(a_v(p00), a_c(p00), a_v(p10)) + 1) >> 1;
    } else {
        const anc_t p01 = e[stride];
        const anc_t p11 = (DX == 3) ? e[stride + 1] : p01;
        v = (lat<DX>(a_v(p00), a_c(p00), a_v(p10)) + lat<DX>(a_i(p01), a_h(p01), a_i(p11)) + 1) >> 1;
    }
    if (DX == 0 && DY == 0) return v;
    return v < 0 ? 0 : (v > 255 ? 255 : v);
}

// Lane geometry for an N x N block split over WPC wavefronts, held as dword
// units (4 pixels each).  N >= 16: lane l of wavefront w owns units
// (64 w + l) + 64 WPC u, u < U.  N == 8: 16 units, so the wavefront's four
// DPP rows price four candidates (G == 4).  Integer candidates are priced B
// at a time (all their loads issued before the first SAD).
template <int N, int WPC> struct Geo {
    static constexpr int DW = N / 4;
    static constexpr int UNITS = N * DW;
    static constexpr int U = UNITS >= 64 * WPC ? UNITS / (64 * WPC) : 1;
    static constexpr int G = UNITS >= 64 ? 1 : 64 / UNITS;
    static constexpr int B = U <= 2 ? 6 : (U <= 4 ? 3 : 2);
    static constexpr int WIN = N + 2;  // quarter-pel anchor window side
};

template <int N, int WPC>
struct Shared {
    anc_t win[(N + 2) * (N + 2)];  // quarter-pel anchor window
    int red[2][WPC][8];            // cross-wavefront partial sums (double-buffered)
};

// Reference luma window staged in LDS by a tile workgroup: samples
// [x0,x1) x [y0,y1) of the reference, row r at byte r * kWinPitch.
constexpr int kWinMargin = 40;                  // samples around the 64x64 centre tile
constexpr int kWinSide = 64 + 2 * kWinMargin;   // 144
constexpr int kWinPitch = kWinSide + 4;         // 148 bytes = 37 dwords per row (odd: banks)
constexpr int kWinWords = kWinSide * kWinPitch / 4;

struct LumaWin {
    const uint32_t *w;  // LDS
    int x0, y0, x1, y1;
};

// Full-pel candidate search + quarter-pel refinement of one call by WPC
// wavefronts.  qw ((N+2)^2 anchors) and red ([2][WPC][8] ints) are LDS areas
// owned by this call's wavefronts.  With WIN, integer candidates that lie
// inside the LDS luma window are read from it instead of the plane.  Without
// QWIN the quarter-pel pass reads the anchor plane directly and qw is unused.
template <int N, int WPC, bool WIN = false, bool QWIN = true>
struct Searcher {
    using Gm = Geo<N, WPC>;
    const PicDev &R;
    anc_t *qw;
    int *red;
    LumaWin lw;
    uint32_t o[Gm::U];    // current block, packed bytes per unit
    int urow[Gm::U], ucol[Gm::U];
    int lane, wave;
    int tid;              // rank among the call's 64 * WPC threads
    int grp;              // candidate group (N == 8)
    int o8;               // N == 8: this lane's current-block pixel
    int phase;            // which half of red the next exchange uses
    gptr32 cplane;        // current picture
    size_t cblk;          // byte offset of the current block in it
    int cpitch;
    gptr32 rplane;        // reference picture
    gptr64 aplane;        // reference anchors
    Counts cnt;

    __device__ Searcher(const PicDev &ref, const PicDev &cur, int px, int py, anc_t *qwin,
                        int *redbuf, LumaWin win = LumaWin{nullptr, 0, 0, 0, 0})
        : R(ref), qw(qwin), red(redbuf), lw(win), phase(0), cplane(as_global32(cur.y)),
          cblk(static_cast<size_t>(py) * cur.pitch + px), cpitch(cur.pitch),
          rplane(as_global32(ref.y)), aplane(as_global64(ref.an))
    {
        // An empty asm makes the lane index opaque, so the lane-derived
        // offsets of each search stay inside the loop that runs it instead of
        // being hoisted and held live across every search of a kernel (which
        // cost the predictor chain ~100 registers).
        int t = static_cast<int>(threadIdx.x);
        asm volatile("" : "+v"(t));
        lane = t & 63;
        tid = t & (64 * WPC - 1);
        wave = WPC == 1 ? 0 : __builtin_amdgcn_readfirstlane(tid >> 6);
        grp = Gm::G == 1 ? 0 : lane / Gm::UNITS;
#pragma unroll
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = Gm::G == 1 ? wave * 64 + lane + 64 * WPC * u : lane % Gm::UNITS;
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

    // Whole-block sums of K wave-uniform partials: with several wavefronts
    // per block the partials meet in LDS (one barrier); every wavefront gets
    // the same totals, so all take the same walk.
    template <int K>
    __device__ __forceinline__ void block_sum(int32_t (&v)[K])
    {
        if (WPC == 1) return;
        if (lane == 0) {
#pragma unroll
            for (int k = 0; k < K; ++k) red[(phase * WPC + wave) * 8 + k] = v[k];
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; ++k) {
            int t = 0;
#pragma unroll
            for (int w = 0; w < WPC; ++w) t += red[(phase * WPC + w) * 8 + k];
            v[k] = __builtin_amdgcn_readfirstlane(t);
        }
        phase ^= 1;
    }

    // Whether the N x N block at (x,y) lies inside the LDS luma window.
    __device__ __forceinline__ bool inwin(int x, int y) const
    {
        return WIN && x >= lw.x0 && y >= lw.y0 && x + N <= lw.x1 && y + N <= lw.y1;
    }

    // Four reference samples of unit u of the block at (x,y): from the LDS
    // window when `w` (the block lies inside it), else from the plane.
    // Makes this call's LDS writes visible to all of its threads.  A call of
    // one wavefront syncs only that wavefront (a tile workgroup's wavefronts
    // work on different calls); LDS operations of one wavefront complete in
    // order, so ordering the compiler is enough.
    __device__ __forceinline__ void sync_call() const
    {
        if (WPC == 1) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        } else {
            __syncthreads();
        }
    }

    __device__ __forceinline__ uint32_t ref4(int x, int y, int u, bool w = false) const
    {
        if (WIN && w) {
            const int a = (y - lw.y0 + urow[u]) * kWinPitch + (x - lw.x0) + ucol[u];
            const uint32_t *q = lw.w + (a >> 2);
            return __builtin_amdgcn_alignbyte(q[1], q[0], static_cast<uint32_t>(a & 3));
        }
        return ld4g(rplane, static_cast<size_t>(y + urow[u]) * R.pitch + x + ucol[u]);
    }

    // SAD of the block at full-pel (x,y), which must fit in the frame.
    __device__ __forceinline__ int32_t sad1(int x, int y)
    {
        if (Gm::G == 1) {
            uint32_t acc = 0;
            const bool w = inwin(x, y);
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) acc = __builtin_amdgcn_sad_u8(ref4(x, y, u, w), o[u], acc);
            int32_t v[1] = {wave_sum(static_cast<int>(acc))};
            block_sum<1>(v);
            return v[0];
        }
        const int v = row_scan(static_cast<int>(__builtin_amdgcn_sad_u8(ref4(x, y, 0, inwin(x, y)), o[0], 0u)));
        return __builtin_amdgcn_readlane(v, 15);  // group 0 priced it; all groups did
    }

    // N >= 16: prices candidates first .. first+B-1 of pattern P (those < K)
    // around (bx,by), loads first; off-frame ones get kInf.
    template <int P>
    __device__ __forceinline__ void price_batch(int bx, int by, int scale, int first,
                                                int32_t (&s)[Gm::B])
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
            const bool w = inwin(lx, ly);
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) v[b][u] = ref4(lx, ly, u, w);
        }
#pragma unroll
        for (int b = 0; b < Gm::B; ++b) {
            uint32_t acc = 0;
#pragma unroll
            for (int u = 0; u < Gm::U; ++u) acc = __builtin_amdgcn_sad_u8(v[b][u], o[u], acc);
            s[b] = wave_sum(static_cast<int>(acc));
        }
        block_sum<Gm::B>(s);
#pragma unroll
        for (int b = 0; b < Gm::B; ++b)
            if (!ok[b]) s[b] = kInf;
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
            // one source for the whole wavefront: the window only when all
            // four groups' candidates lie inside it
            const bool w = WIN && __ballot(!inwin(x, y)) == 0;
            const int v = row_scan(static_cast<int>(__builtin_amdgcn_sad_u8(ref4(x, y, 0, w), o[0], 0u)));
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

    // SAD of a quarter-pel candidate read directly from the planes (edge path),
    // with the candidate's own class (1 six-tap, 2 bilinear).
    __device__ __forceinline__ int32_t qsad_direct(int qx, int qy, int cls)
    {
        const int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
        const gptr64 an = aplane;
        if (N == 8) {
            const int x = fx + (lane & 7), y = fy + (lane >> 3);
            int v;
            if (cls == 1) v = qpix_any(an + static_cast<size_t>(y) * R.apitch + x, R.apitch, dx, dy);
            else v = bil_px(R, x, y, dx, dy);
            const int d = o8 - v;
            return wave_sum(d < 0 ? -d : d);
        }
        // a rolled loop (this is the rare edge path): unit coordinates are
        // recomputed and the current block re-read, so no register array is
        // indexed dynamically
        uint32_t acc = 0;
#pragma unroll 1
        for (int u = 0; u < Gm::U; ++u) {
            const int unit = wave * 64 + lane + 64 * WPC * u;
            const int ur = unit / Gm::DW, uc = (unit % Gm::DW) * 4;
            const int x = fx + uc, y = fy + ur;
            uint32_t pk = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                int v;
                if (cls == 1) v = qpix_any(an + static_cast<size_t>(y) * R.apitch + x + j, R.apitch, dx, dy);
                else v = bil_px(R, x + j, y, dx, dy);
                pk |= static_cast<uint32_t>(v) << (8 * j);
            }
            acc = __builtin_amdgcn_sad_u8(pk, ld4g(cplane, cblk + static_cast<size_t>(ur) * cpitch + uc), acc);
        }
        int32_t v[1] = {wave_sum(static_cast<int>(acc))};
        block_sum<1>(v);
        return v[0];
    }

    template <typename E>
    __device__ static __forceinline__ int qpix_any(E e, int stride, int dx, int dy)
    {
        switch ((dy << 2) | dx) {
        case 0: return qpix<0, 0>(e, stride);
        case 1: return qpix<1, 0>(e, stride);
        case 2: return qpix<2, 0>(e, stride);
        case 3: return qpix<3, 0>(e, stride);
        case 4: return qpix<0, 1>(e, stride);
        case 5: return qpix<1, 1>(e, stride);
        case 6: return qpix<2, 1>(e, stride);
        case 7: return qpix<3, 1>(e, stride);
        case 8: return qpix<0, 2>(e, stride);
        case 9: return qpix<1, 2>(e, stride);
        case 10: return qpix<2, 2>(e, stride);
        case 11: return qpix<3, 2>(e, stride);
        case 12: return qpix<0, 3>(e, stride);
        case 13: return qpix<1, 3>(e, stride);
        case 14: return qpix<2, 3>(e, stride);
        default: return qpix<3, 3>(e, stride);
        }
    }

    // SAD of a quarter-pel candidate from the LDS window: ox, oy (0 or 1)
    // locate its sample origin in the window.
    template <int DX, int DY>
    __device__ __forceinline__ int32_t qsad_win(int ox, int oy)
    {
        const anc_t *w = qw + oy * Gm::WIN + ox;
        if (N == 8) {
            const int d = o8 - qpix<DX, DY>(w + (lane >> 3) * Gm::WIN + (lane & 7), Gm::WIN);
            return wave_sum(d < 0 ? -d : d);
        }
        uint32_t acc = 0;
#pragma unroll
        for (int u = 0; u < Gm::U; ++u) {
            const anc_t *e = w + urow[u] * Gm::WIN + ucol[u];
            const uint32_t pk = static_cast<uint32_t>(qpix<DX, DY>(e, Gm::WIN)) |
                                (static_cast<uint32_t>(qpix<DX, DY>(e + 1, Gm::WIN)) << 8) |
                                (static_cast<uint32_t>(qpix<DX, DY>(e + 2, Gm::WIN)) << 16) |
                                (static_cast<uint32_t>(qpix<DX, DY>(e + 3, Gm::WIN)) << 24);
            acc = __builtin_amdgcn_sad_u8(pk, o[u], acc);
            if (Gm::U >= 8) __builtin_amdgcn_sched_barrier(0);
        }
        int32_t v[1] = {wave_sum(static_cast<int>(acc))};
        block_sum<1>(v);
        return v[0];
    }

    __device__ __forceinline__ int32_t qsad_win_any(int ox, int oy, int dx, int dy)
    {
        switch ((dy << 2) | dx) {
        case 0: return qsad_win<0, 0>(ox, oy);
        case 1: return qsad_win<1, 0>(ox, oy);
        case 2: return qsad_win<2, 0>(ox, oy);
        case 3: return qsad_win<3, 0>(ox, oy);
        case 4: return qsad_win<0, 1>(ox, oy);
        case 5: return qsad_win<1, 1>(ox, oy);
        case 6: return qsad_win<2, 1>(ox, oy);
        case 7: return qsad_win<3, 1>(ox, oy);
        case 8: return qsad_win<0, 2>(ox, oy);
        case 9: return qsad_win<1, 2>(ox, oy);
        case 10: return qsad_win<2, 2>(ox, oy);
        case 11: return qsad_win<3, 2>(ox, oy);
        case 12: return qsad_win<0, 3>(ox, oy);
        case 13: return qsad_win<1, 3>(ox, oy);
        case 14: return qsad_win<2, 3>(ox, oy);
        default: return qsad_win<3, 3>(ox, oy);
        }
    }

    // SAD of a bilinear quarter-pel candidate (reference src/motion.c:111-116,
    // 330-336) from the sample fields of the LDS window: ox, oy (0 or 1)
    // locate its sample origin in the window, dx, dy its fraction.
    __device__ __forceinline__ int32_t qsad_win_bil(int ox, int oy, int dx, int dy)
    {
        const anc_t *w = qw + oy * Gm::WIN + ox;
        const auto px = [dx, dy](const anc_t *e) {
            const int top = (a_i(e[0]) * (4 - dx) + a_i(e[1]) * dx + 2) >> 2;
            const int bot = (a_i(e[Gm::WIN]) * (4 - dx) + a_i(e[Gm::WIN + 1]) * dx + 2) >> 2;
            return (top * (4 - dy) + bot * dy + 2) >> 2;
        };
        if (N == 8) {
            const int d = o8 - px(w + (lane >> 3) * Gm::WIN + (lane & 7));
            return wave_sum(d < 0 ? -d : d);
        }
        uint32_t acc = 0;
#pragma unroll
        for (int u = 0; u < Gm::U; ++u) {
            const anc_t *e = w + urow[u] * Gm::WIN + ucol[u];
            const uint32_t pk = static_cast<uint32_t>(px(e)) | (static_cast<uint32_t>(px(e + 1)) << 8) |
                                (static_cast<uint32_t>(px(e + 2)) << 16) |
                                (static_cast<uint32_t>(px(e + 3)) << 24);
            acc = __builtin_amdgcn_sad_u8(pk, o[u], acc);
        }
        int32_t v[1] = {wave_sum(static_cast<int>(acc))};
        block_sum<1>(v);
        return v[0];
    }

    // Quarter-pel refinement from the planes, every candidate priced with its
    // own class (six-tap or bilinear) straight from the anchor and luma planes.
    __device__ __forceinline__ void qpel_direct(int bx, int by, int &qx, int &qy, int32_t &best)
    {
        constexpr int K = Pat<PAT_QPEL>::K;
        qx = bx * 4;
        qy = by * 4;
#pragma unroll 1
        for (int k = 0; k < K; ++k) {
            const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
            const int cls = qclass(R, N, tx, ty);
            if (!cls) continue;
            if (cls == 1) cnt.q6++; else cnt.qb++;
            const int32_t sv = qsad_direct(tx, ty, cls);
            if (sv < best) { best = sv; qx = tx; qy = ty; }
        }
    }

    // Quarter-pel refinement, one greedy pass in the reference's order
    // (reference src/motion.c:450-481).  (bx,by) is the integer best.  Every
    // sample position the pass can reach lies in fx in {bx-1, bx}, fy in
    // {by-1, by}, so the block's (N+2)^2 anchor window at (bx-1, by-1) is
    // staged in LDS once (coordinates clamped into the picture: near an edge
    // the clamped entries are never read) and each candidate is priced from
    // it with its own class: six-tap from the half samples, bilinear from
    // the sample fields, or skipped (reference src/motion.c:455-464).
    __device__ __forceinline__ void qpel(int bx, int by, int &qx, int &qy, int32_t &best)
    {
        constexpr int K = Pat<PAT_QPEL>::K;
        qx = bx * 4;
        qy = by * 4;
        if (!QWIN) {
            qpel_direct(bx, by, qx, qy, best);
            return;
        }
        const bool six = bx - 1 >= 2 && by - 1 >= 2 && bx + N + 2 < R.w && by + N + 2 < R.h;
        const int x0 = bx - 1, y0 = by - 1;  // window origin (sample coordinates)
        constexpr int WN = Gm::WIN * Gm::WIN;
        // up to 8 loads per thread in flight before their stores (a rolled
        // loop waited for each: a 34x34 window is 19 loads per lane, a 66x66
        // one 9 per thread of 8 wavefronts)
        constexpr int kStep = 64 * WPC, kBatch = 8;
        for (int i0 = tid; i0 < WN; i0 += kBatch * kStep) {
            anc_t v[kBatch];
#pragma unroll
            for (int b = 0; b < kBatch; ++b) {
                const int i = i0 + b * kStep;
                const int r = i / Gm::WIN, c = i - r * Gm::WIN;
                const int y = min(max(y0 + r, 0), R.h - 1), x = min(max(x0 + c, 0), R.w - 1);
                if (i < WN) v[b] = aplane[static_cast<size_t>(y) * R.apitch + x];
            }
#pragma unroll
            for (int b = 0; b < kBatch; ++b)
                if (i0 + b * kStep < WN) qw[i0 + b * kStep] = v[b];
        }
        sync_call();
#pragma unroll 1
        for (int k = 0; k < K; ++k) {
            const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
            const int cls = six ? 1 : qclass(R, N, tx, ty);
            if (!cls) continue;
            int32_t sv;
            if (cls == 1) {
                cnt.q6++;
                sv = qsad_win_any((tx >> 2) - x0, (ty >> 2) - y0, tx & 3, ty & 3);
            } else {
                cnt.qb++;
                sv = qsad_win_bil((tx >> 2) - x0, (ty >> 2) - y0, tx & 3, ty & 3);
            }
            if (sv < best) { best = sv; qx = tx; qy = ty; }
        }
    }
};

template <int N, int WPC, bool WIN = false, bool QWIN = true>
__device__ void search(const PicDev &R, const PicDev &C, const tcme_call &c, tcme_result *res,
                       tcme_stats *st, anc_t *qw, int *red,
                       LumaWin lw = LumaWin{nullptr, 0, 0, 0, 0})
{
    Searcher<N, WPC, WIN, QWIN> S(R, C, c.px, c.py, qw, red, lw);
    // Centre clamp with tc_clip's order: low bound first (motion.c:391-392).
    int bx = c.cx < 0 ? 0 : (c.cx > R.w - N ? R.w - N : c.cx);
    int by = c.cy < 0 ? 0 : (c.cy > R.h - N ? R.h - N : c.cy);
    int32_t best;
    if (fits(R, N, bx, by)) { best = S.sad1(bx, by); S.cnt.ie++; }
    else { best = kInf; S.cnt.oob++; }
    int qx, qy;
    if (R.w >= N && R.h >= N) {
        const int scale = search_scale(c.sr, R.w, R.h);
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
    if (S.tid == 0) {
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

__device__ __forceinline__ void put_status(tcme_result *r, tcme_stats *st, int status)
{
    r->mv_x = 0;
    r->mv_y = 0;
    r->sad = kInf;
    r->status = status;
    if (st) { st->int_evals = st->q6_evals = st->qbil_evals = st->int_oob = 0; }
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
    if (!block_fits(c.px, c.py, c.blk, C.w, C.h)) return TCME_E_ARG;
    return 0;
}

// Block index -> call index so that each of the 8 XCDs (which receive
// workgroups round-robin) works through one contiguous run of the call list:
// neighbouring calls read neighbouring reference rows, which then stay in
// that XCD's L2.
__device__ __forceinline__ int xcd_slot(int b, int nb)
{
    constexpr int kXcd = 8;
    const int x = b % kXcd, slot = b / kXcd;
    const int q = nb / kXcd, r = nb % kXcd;
    return x * q + (x < r ? x : r) + slot;
}

// Workgroup -> chunk of the tile kernel.  Workgroup b runs on XCD b % 8; in
// each round of 8K chunks, XCD x takes the K consecutive chunks x*K .. x*K+K-1
// in turn.  At K = 8 (two CTUs against four references, for CTU-major lists)
// a CTU's right neighbour against the same reference is searched on the same
// XCD, whose L2 still holds much of its window, while every XCD works on every
// part of the frame: a contiguous eighth of the list per XCD left the XCDs
// with the most motion working alone at the end of the launch.
__device__ __forceinline__ int tile_slot(int b, int nb)
{
    constexpr int K = 8, RR = 8 * K;
    if (b >= nb / RR * RR) return b;  // the last, partial round in order
    const int t = b % RR;
    return b - t + (t % 8) * K + t / 8;
}

// One workgroup of WPC wavefronts per call, block size N only: each size has
// its own instance (the register and LDS budgets differ ~16x between 8x8 and
// 64x64).  A workgroup whose call has another size exits at once; the 64
// instance also answers descriptors with an unsupported size.
template <int N, int WPC>
__global__ __launch_bounds__(64 * WPC) void me_search_kernel(const PicDev *__restrict__ table,
                                                              int npics,
                                                              const tcme_call *__restrict__ calls,
                                                              int n, tcme_result *__restrict__ res,
                                                              tcme_stats *__restrict__ stats)
{
    __shared__ Shared<N, WPC> sh;
    const int i = xcd_slot(blockIdx.x, gridDim.x);
    if (i >= n) return;
    const tcme_call c = calls[i];
    const bool mine = c.blk == N || (N == 64 && c.blk != 8 && c.blk != 16 && c.blk != 32);
    if (!mine) return;
    PicDev R{}, C{};
    const int status = check_call(table, npics, c, R, C);
    tcme_stats *st = stats ? stats + i : nullptr;
    if (status) {
        if (threadIdx.x == 0) put_status(res + i, st, status);
        return;
    }
    search<N, WPC>(R, C, c, res + i, st, sh.win, &sh.red[0][0][0]);
}

// The per-call drop-in: one call, passed by value; the current block (N x N
// bytes, row pitch N) is read from mapped host memory and the result written
// to mapped host memory, status last, so a call costs one launch and no
// copies, and the host can poll the status instead of synchronising.  The
// host has checked the block size and that `ref_pic` holds a reference
// picture.
template <int N, int WPC>
__global__ __launch_bounds__(64 * WPC) void me_dropin_kernel(const PicDev *__restrict__ table,
                                                              tcme_call c,
                                                              const uint8_t *__restrict__ blk,
                                                              tcme_result *__restrict__ res)
{
    __shared__ Shared<N, WPC> sh;
    __shared__ tcme_result r;
    const PicDev R = table[c.ref_pic];
    const PicDev C{blk, nullptr, N, N, N, 0};
    c.px = 0;
    c.py = 0;
    search<N, WPC>(R, C, c, &r, nullptr, sh.win, &sh.red[0][0][0]);
    __syncthreads();
    // the host polls the status: it goes last, after the other fields are
    // visible system-wide
    if (threadIdx.x == 0) {
        res->mv_x = r.mv_x;
        res->mv_y = r.mv_y;
        res->sad = r.sad;
        __threadfence_system();
        res->status = r.status;
    }
}

// The drop-in's single call for a block size other than 8, 16, 32 or 64 (the
// reference takes any n, src/motion.c:384-392; its encoders use those four).
// One workgroup of 256 threads walks the reference's sequential search with
// wave-uniform decisions; each candidate's SAD is a block-wide sum (each
// thread a strided share of the n x n samples, then LDS).  The current block
// (pitch n, mapped host memory) is staged in LDS when n <= 64 and read
// through otherwise.  Result as me_dropin_kernel: status last, after a
// system-scope fence.
constexpr int kGenericLds = 64 * 64;

struct Generic {
    const PicDev &R;
    const uint8_t *blk;      // LDS copy (n <= 64) or the host block
    gptr8 hblk;
    bool lds;
    int n, tid;
    int *part;               // 4 ints of LDS
    Counts cnt;

    __device__ int cur(int i) const { return lds ? blk[i] : hblk[i]; }

    __device__ int32_t sum(int v)
    {
        const int wave = tid >> 6;
        v = wave_sum(v);
        __syncthreads();
        if ((tid & 63) == 0) part[wave] = v;
        __syncthreads();
        return part[0] + part[1] + part[2] + part[3];
    }

    __device__ int32_t sad(int x, int y)
    {
        const gptr8 p = as_global8(R.y);
        int acc = 0;
        for (int i = tid; i < n * n; i += 256) {
            const int r = i / n, c = i - r * n;
            const int d = cur(i) - p[static_cast<size_t>(y + r) * R.pitch + x + c];
            acc += d < 0 ? -d : d;
        }
        return sum(acc);
    }

    __device__ int32_t qsad(int qx, int qy, int cls)
    {
        const int fx = qx >> 2, fy = qy >> 2, dx = qx & 3, dy = qy & 3;
        const gptr64 an = as_global64(R.an);
        int acc = 0;
        for (int i = tid; i < n * n; i += 256) {
            const int r = i / n, c = i - r * n;
            const int v = cls == 1
                ? Searcher<8, 1>::qpix_any(an + static_cast<size_t>(fy + r) * R.apitch + fx + c, R.apitch, dx, dy)
                : bil_px(R, fx + c, fy + r, dx, dy);
            const int d = cur(i) - v;
            acc += d < 0 ? -d : d;
        }
        return sum(acc);
    }

    template <int P>
    __device__ void stage(int scale, int passes, int &bx, int &by, int32_t &best)
    {
        for (int it = 0; it < passes; ++it) {
            bool moved = false;
            for (int k = 0; k < Pat<P>::K; ++k) {
                const int x = bx + Pat<P>::x(k) * scale, y = by + Pat<P>::y(k) * scale;
                if (!fits(R, n, x, y)) { cnt.oob++; continue; }
                cnt.ie++;
                const int32_t s = sad(x, y);
                if (s < best) { best = s; bx = x; by = y; moved = true; }
            }
            if (!moved) break;
        }
    }
};

__global__ __launch_bounds__(256) void me_generic_kernel(const PicDev *__restrict__ table, tcme_call c,
                                                         const uint8_t *__restrict__ blk,
                                                         tcme_result *__restrict__ res)
{
    __shared__ uint8_t sblk[kGenericLds];
    __shared__ int part[4];
    const PicDev R = table[c.ref_pic];
    const int n = c.blk, tid = static_cast<int>(threadIdx.x);
    const bool lds = n <= 64;
    if (lds) {
        for (int i = tid; i < n * n; i += 256) sblk[i] = as_global8(blk)[i];
        __syncthreads();
    }
    Generic G{R, sblk, as_global8(blk), lds, n, tid, part, Counts{}};
    // centre clamp with tc_clip's order (src/motion.c:391-392)
    int bx = c.cx < 0 ? 0 : (c.cx > R.w - n ? R.w - n : c.cx);
    int by = c.cy < 0 ? 0 : (c.cy > R.h - n ? R.h - n : c.cy);
    int32_t best = fits(R, n, bx, by) ? G.sad(bx, by) : kInf;
    int qx = bx * 4, qy = by * 4;
    if (R.w >= n && R.h >= n) {
        G.template stage<PAT_LARGE>(search_scale(c.sr, R.w, R.h), 2, bx, by, best);
        G.template stage<PAT_SMALL>(1, 3, bx, by, best);
        G.template stage<PAT_DIAMOND>(1, 4, bx, by, best);
        qx = bx * 4;
        qy = by * 4;
        for (int k = 0; k < Pat<PAT_QPEL>::K; ++k) {
            const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
            const int cls = qclass(R, n, tx, ty);
            if (!cls) continue;
            const int32_t sv = G.qsad(tx, ty, cls);
            if (sv < best) { best = sv; qx = tx; qy = ty; }
        }
    }
    if (tid == 0) {
        res->mv_x = qx;
        res->mv_y = qy;
        res->sad = best;
        __threadfence_system();
        res->status = 0;
    }
}

// A short list of calls on resident pictures, one workgroup of 4 wavefronts
// per call, any block size: the frame-level drop-in's searches of a call
// that missed its cache, together with its CTU's other nodes at the same
// displacement.  Calls are read from, and results written to, mapped host
// memory, each result's status last after a system-scope fence, so the host
// can take any one result as soon as its status changes (the others may
// still be running).  The host has validated every call (reference pictures
// with anchors, blocks inside the current picture, sizes 8/16/32/64).  An
// 8x8 or 16x16 call is searched by the first wavefront alone (the others
// leave at once; that search synchronises within its wavefront only).
__global__ __launch_bounds__(256) void me_list_kernel(const PicDev *__restrict__ table,
                                                      const tcme_call *__restrict__ calls,
                                                      tcme_result *__restrict__ res)
{
    __shared__ union {
        Shared<8, 1> s8;
        Shared<16, 1> s16;
        Shared<32, 4> s32;
        Shared<64, 4> s64;
    } sh;
    __shared__ tcme_result r;
    const tcme_call c = calls[blockIdx.x];
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    const PicDev R = table[c.ref_pic], C = table[c.cur_pic];
    switch (c.blk) {
    case 8:
        if (wave) return;
        search<8, 1>(R, C, c, &r, nullptr, sh.s8.win, &sh.s8.red[0][0][0]);
        break;
    case 16:
        if (wave) return;
        search<16, 1>(R, C, c, &r, nullptr, sh.s16.win, &sh.s16.red[0][0][0]);
        break;
    case 32:
        search<32, 4>(R, C, c, &r, nullptr, sh.s32.win, &sh.s32.red[0][0][0]);
        break;
    default:
        search<64, 4>(R, C, c, &r, nullptr, sh.s64.win, &sh.s64.red[0][0][0]);
        break;
    }
    // thread 0 wrote r itself (it is the call's rank-0 thread in every case)
    if (threadIdx.x == 0) {
        tcme_result *o = res + blockIdx.x;
        o->mv_x = r.mv_x;
        o->mv_y = r.mv_y;
        o->sad = r.sad;
        __threadfence_system();
        o->status = r.status;
    }
}

// ---- ME-only predictor chain ------------------------------------------------
// One wavefront per CTU walks the CTU's quadtree in the
// reference encoder's decision order (src/encoder.c:1292-1397: a node as a
// leaf, its four children as leaves, and, when the children's cost is not
// above the parent's, each child again as a subtree; nodes not wholly in the
// frame are split without a search).  Every leaf is centred at qt_mvp
// (src/encoder.c:405-427) of the CTU's 8x8 displacement grid, reset to
// "unavailable" per CTU (:1425-1428); it searches each reference in turn and
// keeps the cheapest J = SAD + lambda * (se_bits(mvd.x) + se_bits(mvd.y))
// (first on ties), and its cells take the chosen displacement.  The control
// flow is uniform: every lane holds the same walk state; the grid lives in
// LDS.  A CTU's searches run one after another, so wider workgroups would
// mostly wait; one wavefront keeps the LDS per CTU small (the 64x64
// quarter-pel pass reads the anchor plane directly) and many CTUs resident.
struct ChainJob {
    int32_t cur_pic, nrefs, ref_pic[4], sr, lambda, first, cols;
};
struct ChainArgs {
    int32_t njobs, pad;
    ChainJob job[kChainFrames];
};

// One level of the quadtree walk: the node's first cell, its stage, the
// next child, and the costs of the node and of its children as leaves.
struct ChainLevel {
    int32_t cx, cy, stage, q;
    int64_t parent, child;
};

struct ChainLds {
    union {
        Shared<8, 1> s8;
        Shared<16, 1> s16;
        Shared<32, 1> s32;
    } sh;
    tcme_result r;
    int16_t dx[64], dy[64];
    uint8_t un[64];                          // 1: unavailable (intra)
    int16_t sdx[2][3][64], sdy[2][3][64];    // saved regions per depth: pre, parent
    uint8_t sun[2][3][64];
    ChainLevel lv[4];                        // walk stack, one level per depth
};

struct ChainWalk {
    const PicDev *table;
    PicDev C;
    ChainJob j;
    int ctu_x, ctu_y, w, h;
    tcme_chain_call *out;
    int nout;
};

__device__ __forceinline__ int se_bits(int v)
{
    const uint32_t m = v > 0 ? 2u * static_cast<uint32_t>(v) - 1u : (v < 0 ? 2u * static_cast<uint32_t>(-v) : 0u);
    return 2 * (32 - __builtin_clz(m + 1u)) - 1;  // m + 1 >= 1
}

// qt_mvp (src/encoder.c:405-427) of cell (cx, cy) from the LDS grid.
__device__ __forceinline__ void chain_mvp(const ChainLds &L, int cx, int cy, int &mx, int &my)
{
    int ax = 0, ay = 0, bx = 0, by = 0, qx = 0, qy = 0;
    bool ha = false, hb = false, hc = false;
    if (cx > 0 && !L.un[cy * 8 + cx - 1]) { ax = L.dx[cy * 8 + cx - 1]; ay = L.dy[cy * 8 + cx - 1]; ha = true; }
    if (cy > 0 && !L.un[(cy - 1) * 8 + cx]) { bx = L.dx[(cy - 1) * 8 + cx]; by = L.dy[(cy - 1) * 8 + cx]; hb = true; }
    if (cy > 0 && cx < 7 && !L.un[(cy - 1) * 8 + cx + 1]) {
        qx = L.dx[(cy - 1) * 8 + cx + 1]; qy = L.dy[(cy - 1) * 8 + cx + 1]; hc = true;
    }
    if (!ha && !hb && !hc) { mx = 0; my = 0; return; }
    if (!ha) { ax = hb ? bx : qx; ay = hb ? by : qy; }
    if (!hb) { bx = ha ? ax : qx; by = ha ? ay : qy; }
    if (!hc) { qx = ha ? ax : bx; qy = ha ? ay : by; }
    const int64_t ma = static_cast<int64_t>(ax) * ax + static_cast<int64_t>(ay) * ay;
    const int64_t mb = static_cast<int64_t>(bx) * bx + static_cast<int64_t>(by) * by;
    const int64_t mc = static_cast<int64_t>(qx) * qx + static_cast<int64_t>(qy) * qy;
    if (mb < ma && mb <= mc) { mx = bx; my = by; }
    else if (mc < ma && mc < mb) { mx = qx; my = qy; }
    else { mx = ax; my = ay; }
}

// Copies the cs x cs region at (cx, cy) of the grid to or from save slot
// (k, d); one cell per thread of the first 64.
__device__ __forceinline__ void chain_copy(ChainLds &L, int cx, int cy, int cs, int k, int d, bool save)
{
    const int t = static_cast<int>(threadIdx.x);
    if (t < cs * cs) {
        const int i = (cy + t / cs) * 8 + cx + t % cs;
        if (save) { L.sdx[k][d][t] = L.dx[i]; L.sdy[k][d][t] = L.dy[i]; L.sun[k][d][t] = L.un[i]; }
        else { L.dx[i] = L.sdx[k][d][t]; L.dy[i] = L.sdy[k][d][t]; L.un[i] = L.sun[k][d][t]; }
    }
    __syncthreads();
}

template <int D>
__device__ __forceinline__ int64_t chain_leaf(ChainLds &L, ChainWalk &W, int cx, int cy)
{
    constexpr int N = 64 >> D;
    const int px = W.ctu_x + 8 * cx, py = W.ctu_y + 8 * cy;
    int mx, my;
    chain_mvp(L, cx, cy, mx, my);
    const int ccx = (mx + 4 * px) >> 2, ccy = (my + 4 * py) >> 2;
    int64_t best = INT64_MAX;
    int bdx = 0, bdy = 0;
    for (int r = 0; r < W.j.nrefs; ++r) {
        // a select, not an index: indexing the walk state by r would keep it
        // in private memory
        const int rp = r == 0 ? W.j.ref_pic[0] : r == 1 ? W.j.ref_pic[1] : r == 2 ? W.j.ref_pic[2] : W.j.ref_pic[3];
        const PicDev R = W.table[rp];
        const tcme_call c{rp, W.j.cur_pic, px, py, ccx, ccy, N, W.j.sr};
        if constexpr (N == 8) search<8, 1>(R, W.C, c, &L.r, nullptr, L.sh.s8.win, &L.sh.s8.red[0][0][0]);
        else if constexpr (N == 16) search<16, 1>(R, W.C, c, &L.r, nullptr, L.sh.s16.win, &L.sh.s16.red[0][0][0]);
        else if constexpr (N == 32) search<32, 1>(R, W.C, c, &L.r, nullptr, L.sh.s32.win, &L.sh.s32.red[0][0][0]);
        else search<64, 1, false, false>(R, W.C, c, &L.r, nullptr, nullptr, L.sh.s8.red[0][0]);
        __syncthreads();
        const int mvx = L.r.mv_x, mvy = L.r.mv_y, sad = L.r.sad;
        __syncthreads();  // r and the search's LDS are reused by the next search
        const int64_t jc = static_cast<int64_t>(sad) +
                           static_cast<int64_t>(W.j.lambda) * (se_bits(mvx - (mx + 4 * px)) + se_bits(mvy - (my + 4 * py)));
        if (threadIdx.x == 0 && W.nout < kChainMaxCalls)
            W.out[W.nout] = tcme_chain_call{r, px, py, ccx, ccy, N, mvx, mvy, sad, 0};
        ++W.nout;
        if (jc < best) { best = jc; bdx = mvx - 4 * px; bdy = mvy - 4 * py; }
    }
    constexpr int cs = N / 8;
    const int t = static_cast<int>(threadIdx.x);
    if (t < cs * cs) {
        const int i = (cy + t / cs) * 8 + cx + t % cs;
        L.dx[i] = static_cast<int16_t>(bdx);
        L.dy[i] = static_cast<int16_t>(bdy);
        L.un[i] = 0;
    }
    __syncthreads();
    return best;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ __forceinline__ int64_t uni64(int64_t v)
{
    const uint64_t u = static_cast<uint64_t>(v);
    const uint32_t hi = static_cast<uint32_t>(uni(static_cast<int>(u >> 32)));
    const uint32_t lo = static_cast<uint32_t>(uni(static_cast<int>(u)));
    return static_cast<int64_t>((static_cast<uint64_t>(hi) << 32) | lo);
}

// Walks the CTU's quadtree with an explicit stack of levels in LDS (every
// lane keeps the same copy), so that each block size's search appears once
// in the code.  A level stands at one of these stages:
enum : int32_t {
    kEnter,        // node just reached
    kPop,          // 8x8 leaf searched: back to the parent
    kOutside,      // node not wholly in the frame: children as subtrees
    kAfterParent,  // node searched as a leaf: now its children as leaves
    kChildLeaves,  // children as leaves, q of them requested
    kSubtrees,     // children again, as subtrees
};

__device__ __forceinline__ void chain_walk(ChainLds &L, ChainWalk &W)
{
    int d = 0;
    L.lv[0] = ChainLevel{0, 0, kEnter, 0, 0, 0};
    int64_t cost = 0;  // of the last leaf searched
    for (;;) {
        int leaf = -1, lcx = 0, lcy = 0;
        while (leaf < 0 && d >= 0) {
            ChainLevel &v = L.lv[d];
            const int cx = uni(v.cx), cy = uni(v.cy), stage = uni(v.stage), q = uni(v.q);
            const int n = 64 >> d, half = n / 16, cs = n / 8;
            const int px = W.ctu_x + 8 * cx, py = W.ctu_y + 8 * cy;
            const bool in_frame = px + n <= W.w && py + n <= W.h;
            switch (stage) {
            case kEnter:
                if (d == 3) {
                    if (in_frame) { leaf = 3; lcx = cx; lcy = cy; v.stage = kPop; }
                    else --d;
                } else if (!in_frame) {
                    v.stage = kOutside;
                    v.q = 0;
                } else {
                    chain_copy(L, cx, cy, cs, 0, d, true);
                    leaf = d; lcx = cx; lcy = cy;
                    v.stage = kAfterParent;
                }
                break;
            case kPop:
                --d;
                break;
            case kOutside:
            case kSubtrees:
                if (q < 4) {
                    v.q = q + 1;
                    L.lv[d + 1] = ChainLevel{cx + (q & 1) * half, cy + (q >> 1) * half, kEnter, 0, 0, 0};
                    ++d;
                } else {
                    --d;
                }
                break;
            case kAfterParent:
                v.parent = cost;
                chain_copy(L, cx, cy, cs, 1, d, true);
                chain_copy(L, cx, cy, cs, 0, d, false);
                v.child = 0;
                v.q = 1;
                v.stage = kChildLeaves;
                leaf = d + 1; lcx = cx; lcy = cy;
                break;
            default: {  // kChildLeaves
                const int64_t child = uni64(v.child) + cost;
                v.child = child;
                if (q < 4) {
                    v.q = q + 1;
                    leaf = d + 1; lcx = cx + (q & 1) * half; lcy = cy + (q >> 1) * half;
                } else if (child <= uni64(v.parent)) {
                    chain_copy(L, cx, cy, cs, 0, d, false);
                    v.stage = kSubtrees;
                    v.q = 0;
                } else {
                    chain_copy(L, cx, cy, cs, 1, d, false);
                    --d;
                }
                break;
            }
            }
        }
        if (leaf < 0) return;
        switch (leaf) {
        case 0: cost = chain_leaf<0>(L, W, lcx, lcy); break;
        case 1: cost = chain_leaf<1>(L, W, lcx, lcy); break;
        case 2: cost = chain_leaf<2>(L, W, lcx, lcy); break;
        default: cost = chain_leaf<3>(L, W, lcx, lcy); break;
        }
    }
}

__global__ __launch_bounds__(64) void me_chain_kernel(const PicDev *__restrict__ table, ChainArgs a,
                                                       tcme_chain_call *__restrict__ calls,
                                                       int32_t *__restrict__ counts)
{
    __shared__ ChainLds L;
    const int b = static_cast<int>(blockIdx.x);
    int k = 0;
    while (k + 1 < a.njobs && b >= a.job[k + 1].first) ++k;
    ChainWalk W;
    W.table = table;
    W.j = a.job[k];
    W.C = table[W.j.cur_pic];
    W.w = W.C.w;
    W.h = W.C.h;
    const int ctu = b - W.j.first;
    W.ctu_x = (ctu % W.j.cols) * 64;
    W.ctu_y = (ctu / W.j.cols) * 64;
    W.out = calls + static_cast<size_t>(b) * kChainMaxCalls;
    W.nout = 0;
    const int t = static_cast<int>(threadIdx.x);
    if (t < 64) { L.dx[t] = 0; L.dy[t] = 0; L.un[t] = 1; }
    __syncthreads();
    chain_walk(L, W);
    if (t == 0) counts[b] = W.nout;
}

// Sum over each group of L consecutive lanes (L = 1, 2, 4, ..., 64, groups
// aligned to L); every lane of a group gets its group's total.  DPP
// butterflies: quad_perm [1,0,3,2], quad_perm [2,3,0,1], row_half_mirror,
// row_mirror, then lane permutes across rows.  A group's lanes are all active
// or all inactive.
template <int L>
__device__ __forceinline__ int group_sum(int v)
{
    if (L >= 2) v += __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    if (L >= 4) v += __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    if (L >= 8) v += __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    if (L >= 16) v += __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
    // across rows: ds_bpermute from lane ^ 16, then lane ^ 32 (no LDS memory)
    const int lane = static_cast<int>(threadIdx.x & 63);
    if (L >= 32) v += __builtin_amdgcn_ds_bpermute((lane ^ 16) << 2, v);
    if (L >= 64) v += __builtin_amdgcn_ds_bpermute((lane ^ 32) << 2, v);
    return v;
}

// Half-sample lattice of the reference around a 64x64 tile, staged in LDS
// from the anchor plane: anchors [x0,x1) x [y0,y1) (at most kLatSide per
// side), lattice row 2r holding the low anchor words of anchor row r
// (sample | hh << 16) and row 2r+1 the high words (vv | cc << 16), one dword
// per anchor, kLatPitch dwords per row.  Lattice entry (Y, X), X counting
// 16-bit halves, is then the sample, hh, vv or cc at half-sample position
// (X/2, Y/2) for (Y, X) even/even, even/odd, odd/even, odd/odd.
constexpr int kLatMargin = 8;
constexpr int kLatSide = 64 + 2 * kLatMargin;   // 80 anchors
constexpr int kLatPitch = kLatSide + 1;         // 81 dwords per lattice row (odd: banks)
constexpr int kLatWords = 2 * kLatSide * kLatPitch;

struct LatWin {
    const uint32_t *w;  // LDS
    int x0, y0, x1, y1;
};

// Two signed 16-bit lanes of one dword (v_pk_* integer instructions).
typedef short pk16 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pk16 as_pk16(uint32_t v) { return __builtin_bit_cast(pk16, v); }
__device__ __forceinline__ uint32_t as_u32(pk16 v) { return __builtin_bit_cast(uint32_t, v); }

// The signed 16-bit field of v starting at bit `bit` (0 or 16).  The
// builtin's result type is unsigned: it is converted before any arithmetic,
// or sums of negative fields would wrap and shift logically.
__device__ __forceinline__ int half16(uint32_t v, int bit)
{
    return static_cast<int>(__builtin_amdgcn_sbfe(static_cast<int>(v), bit, 16));
}

// Full-pel search and quarter-pel refinement of 64/L calls of one block size
// at once: each call is walked by a group of L lanes, lane j of the group
// holding rows [j*R, j*R + R) of the current block (R = N/L).  The walks are
// the reference's sequential loops run per lane (reference
// src/motion.c:384-481), so the walk state and every decision live in vector
// registers and the calls of one wavefront diverge freely.  A full-pel
// candidate costs each lane R rows of reference bytes, from the LDS luma
// window when the block lies inside it, else from the plane; a quarter-pel
// candidate costs R rows of lattice words from LDS.
template <int N, int L>
struct GroupWalk {
    static constexpr int R = N / L;
    static constexpr int DW = N / 4;
    const PicDev &Rf;
    gptr32 rplane;
    LumaWin lw;
    uint32_t o[R * DW];
    int j;

    __device__ GroupWalk(const PicDev &ref, LumaWin win, const PicDev &cur, int px, int py, int lane)
        : Rf(ref), rplane(as_global32(ref.y)), lw(win), j(lane % L)
    {
        const gptr32 cp = as_global32(cur.y);
        const size_t base = static_cast<size_t>(py + j * R) * cur.pitch + px;
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
#pragma unroll
            for (int d = 0; d < DW; ++d)
                o[rr * DW + d] = ld4g(cp, base + static_cast<size_t>(rr) * cur.pitch + 4 * d);
    }

    // SAD of the group's block at full-pel (x,y), which must fit in the frame.
    __device__ __forceinline__ int sad(int x, int y) const
    {
        uint32_t acc = 0;
        if (x >= lw.x0 && y >= lw.y0 && x + N <= lw.x1 && y + N <= lw.y1) {
            const int a = (y - lw.y0 + j * R) * kWinPitch + (x - lw.x0);
            const uint32_t sh = static_cast<uint32_t>(a & 3);
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const uint32_t *q = lw.w + ((a >> 2) + rr * (kWinPitch / 4));
                uint32_t t[DW + 1];
#pragma unroll
                for (int d = 0; d <= DW; ++d) t[d] = q[d];
#pragma unroll
                for (int d = 0; d < DW; ++d)
                    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(t[d + 1], t[d], sh),
                                                  o[rr * DW + d], acc);
            }
        } else {
            const size_t b = static_cast<size_t>(y + j * R) * Rf.pitch + x;
            const uint32_t sh = static_cast<uint32_t>(b & 3);
#pragma unroll
            for (int rr = 0; rr < R; ++rr) {
                const gptr32 q = rplane + ((b >> 2) + static_cast<size_t>(rr) * (Rf.pitch >> 2));
                uint32_t t[DW + 1];
#pragma unroll
                for (int d = 0; d <= DW; ++d) t[d] = q[d];
#pragma unroll
                for (int d = 0; d < DW; ++d)
                    acc = __builtin_amdgcn_sad_u8(__builtin_amdgcn_alignbyte(t[d + 1], t[d], sh),
                                                  o[rr * DW + d], acc);
            }
        }
        return group_sum<L>(static_cast<int>(acc));
    }

    // One pattern stage, exactly the reference's loop: candidates in table
    // order around the current best, a strictly better one moves the centre
    // at once; a pass without a move ends the stage.  (All the group's calls
    // step through the candidates together, so a pass costs K candidate
    // prices per wavefront whatever the calls do.)
    template <int P>
    __device__ __forceinline__ void stage(int scale, int passes, int &bx, int &by, int &best,
                                          int &ie, int &oob) const
    {
        for (int it = 0; it < passes; ++it) {
            bool moved = false;
#pragma unroll
            for (int k = 0; k < Pat<P>::K; ++k) {
                const int x = bx + Pat<P>::x(k) * scale, y = by + Pat<P>::y(k) * scale;
                if (fits(Rf, N, x, y)) {
                    ++ie;
                    const int s = sad(x, y);
                    if (s < best) { best = s; bx = x; by = y; moved = true; }
                } else {
                    ++oob;
                }
            }
            if (!moved) break;
        }
    }

    // Whether the quarter-pel pass around the full-pel best (bx,by) can run
    // from the lattice: every candidate six-tap (the reference's class test,
    // src/motion.c:455-464) and every anchor it reads, [bx-1, bx+N] x
    // [by-1, by+N], inside the lattice window.
    __device__ __forceinline__ bool lattice_ok(const LatWin &lt, int bx, int by) const
    {
        return bx - 1 >= 2 && by - 1 >= 2 && bx + N + 2 < Rf.w && by + N + 2 < Rf.h &&
               bx - 1 >= lt.x0 && by - 1 >= lt.y0 && bx + N < lt.x1 && by + N < lt.y1;
    }

    // SAD of the quarter-pel candidate (tx,ty) from the lattice.  Sample
    // (c,r) of the block sits at quarter-pel (tx + 4c, ty + 4r); with
    // X0 = q >> 1 and X1 = (q + 1) >> 1 per axis its value is
    //   clip((avg(M[Y0][X0], M[Y0][X1]) + avg(M[Y1][X0], M[Y1][X1]) + 1) >> 1),
    // avg(a,b) = (a + b + 1) >> 1, which is the six-tap quarter sample of
    // the reference (src/motion.c:193-260) for every phase: X0 == X1 (or
    // Y0 == Y1) on the full- and half-sample phases.
    __device__ __forceinline__ int qsad(const LatWin &lt, int tx, int ty) const
    {
        const int cx = (tx >> 1) - 2 * lt.x0;           // lattice column of X0 at c = 0
        const uint32_t hs = static_cast<uint32_t>(cx & 1) * 16;
        // byte selectors: the low halves of two dwords (M[.][X0] of two
        // samples) and the halves holding M[.][X1] (high when X1 = X0 + 1)
        const uint32_t sel0 = 0x05040100u;
        const uint32_t sel1 = ((tx + 1) >> 1) != (tx >> 1) ? 0x07060302u : sel0;
        const int y0 = (ty >> 1) - 2 * lt.y0 + 2 * j * R;
        const int y1 = ((ty + 1) >> 1) - 2 * lt.y0 + 2 * j * R;
        const pk16 one = {1, 1}, lo = {0, 0}, hi = {255, 255};
        uint32_t acc = 0;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
            const uint32_t *p0 = lt.w + (y0 + 2 * rr) * kLatPitch + (cx >> 1);
            const uint32_t *p1 = lt.w + (y1 + 2 * rr) * kLatPitch + (cx >> 1);
#pragma unroll
            for (int d = 0; d < DW; ++d) {
                uint32_t a[5], b[5];
#pragma unroll
                for (int t = 0; t < 5; ++t) {
                    a[t] = p0[4 * d + t];
                    b[t] = p1[4 * d + t];
                }
                // sample t's two lattice halves from X0 on: M[.][X0] low,
                // M[.][X0 + 1] high
                uint32_t da[4], db[4];
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    da[t] = __builtin_amdgcn_alignbit(a[t + 1], a[t], hs);
                    db[t] = __builtin_amdgcn_alignbit(b[t + 1], b[t], hs);
                }
                // two samples per packed 16-bit operation (the values stay in
                // [-420, 931], so no lane wraps)
                uint32_t v2[2];
#pragma unroll
                for (int h = 0; h < 2; ++h) {
                    const pk16 a0 = as_pk16(__builtin_amdgcn_perm(da[2 * h + 1], da[2 * h], sel0));
                    const pk16 a1 = as_pk16(__builtin_amdgcn_perm(da[2 * h + 1], da[2 * h], sel1));
                    const pk16 b0 = as_pk16(__builtin_amdgcn_perm(db[2 * h + 1], db[2 * h], sel0));
                    const pk16 b1 = as_pk16(__builtin_amdgcn_perm(db[2 * h + 1], db[2 * h], sel1));
                    const pk16 top = (a0 + a1 + one) >> 1;
                    const pk16 bot = (b0 + b1 + one) >> 1;
                    const pk16 v = (top + bot + one) >> 1;
                    v2[h] = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi));
                }
                // the low byte of each of the four 16-bit results
                const uint32_t pk = __builtin_amdgcn_perm(v2[1], v2[0], 0x06040200u);
                acc = __builtin_amdgcn_sad_u8(pk, o[rr * DW + d], acc);
                // keep each group of four samples' loads with their use
                // (bounds the registers in flight; other wavefronts hide
                // the LDS latency)
                __builtin_amdgcn_sched_barrier(0);
            }
        }
        return group_sum<L>(static_cast<int>(acc));
    }

    // The quarter-pel pass from the lattice (lattice_ok must hold): every
    // candidate is six-tap and in frame.
    __device__ __forceinline__ void qpel(const LatWin &lt, int bx, int by, int &qx, int &qy,
                                         int &best, int &q6) const
    {
        qx = bx * 4;
        qy = by * 4;
#pragma unroll
        for (int k = 0; k < Pat<PAT_QPEL>::K; ++k) {
            const int tx = qx + Pat<PAT_QPEL>::x(k), ty = qy + Pat<PAT_QPEL>::y(k);
            ++q6;
            const int s = qsad(lt, tx, ty);
            if (s < best) { best = s; qx = tx; qy = ty; }
        }
    }
};

// Tile workgroups for the 8x8 and 16x16 calls.  A workgroup takes `chunk`
// consecutive calls, keeps those of size 8 and 16 in list order, and splits
// them into runs of consecutive calls that share the reference picture and the
// 64x64 tile holding their centre.  For each run it stages in LDS, once, the
// reference luma around that tile (kWinMargin samples on each side) and the
// half-sample lattice around it (kLatMargin), cuts the run's calls of each
// size into slots of 64/L calls, and its wavefronts take slots: a slot's calls
// are searched side by side (GroupWalk).  Candidates outside the luma window
// are read from the plane.  Calls whose quarter-pel pass leaves the lattice
// (or needs bilinear edge candidates) are deferred: once every slot is done,
// the lattice's LDS holds one anchor window per wavefront and the wavefronts
// refine the deferred calls one at a time (Searcher::qpel).  Any call order
// gives the same results; lists grouped by (CTU, reference), as an encoder
// issues them, make one run per chunk.
// Lanes per call by block size: 64x64 one call per wavefront, 32x32 four,
// 16x16 and 8x8 eight (each lane then holds 64, 64, 32 or 8 samples of its
// block: one row of an 8x8 block, so a CTU's 64 8x8 calls are 8 slots with
// short candidate chains; 16 calls of two rows each per wavefront measured
// 0.5% slower, 32 of four rows 5% slower).
template <int N> struct Lanes {
    static constexpr int L = N == 64 ? 64 : N == 32 ? 32 : 8;
    static constexpr int C = 64 / L;  // calls per slot
};
__host__ __device__ constexpr int size_class(int n) { return n == 64 ? 0 : n == 32 ? 1 : n == 16 ? 2 : 3; }

template <int WAVES>
struct TileShared {
    uint32_t luma[kWinWords];
    union {
        uint32_t lat[kLatWords];     // the run's lattice, while slots run
        anc_t qw[kLatWords / 2];     // then anchor windows for deferred calls
    };
    tcme_call call[kTileChunkMax];   // the chunk's calls, list order
    int idx[kTileChunkMax];          // their list positions
    uint8_t list[4][kTileChunkMax];  // the current run's calls by size (64, 32, 16, 8)
    // calls whose quarter-pel pass leaves the lattice, by size (64, 32, 8/16):
    // call index and full-pel state
    uint8_t def_ci[3][kTileChunkMax];
    int16_t def_x[3][kTileChunkMax], def_y[3][kTileChunkMax];
    int def_sad[3][kTileChunkMax];
    uint8_t def_ie[3][kTileChunkMax], def_oob[3][kTileChunkMax];
    int nlist[4];
    int ndef[3];
    int nlive;
    int next;                        // work counter, 64 per item taken
    int next32;                      // the same, deferred 32x32 calls
};

template <int N, int WAVES>
__device__ __forceinline__ void run_slot(TileShared<WAVES> &sh, const PicDev *__restrict__ table,
                                         const PicDev &R, const LumaWin &lw, const LatWin &lt,
                                         int first, int cnt, int lane,
                                         tcme_result *__restrict__ res,
                                         tcme_stats *__restrict__ stats)
{
    constexpr int L = Lanes<N>::L;
    constexpr int li = size_class(N);
    constexpr int di = N == 64 ? 0 : N == 32 ? 1 : 2;
    const int g = lane / L;
    bool slow = false;  // the quarter-pel pass leaves the lattice: deferred
    int bx = 0, by = 0, best = kInf, ie = 0, oob = 0;
    if (g < cnt) {
        const int ci = sh.list[li][first + g];
        const tcme_call &c = sh.call[ci];
        const PicDev C = table[c.cur_pic];
        GroupWalk<N, L> G(R, lw, C, c.px, c.py, lane);
        // centre clamp with tc_clip's order, low bound first (motion.c:391-392)
        bx = c.cx < 0 ? 0 : (c.cx > R.w - N ? R.w - N : c.cx);
        by = c.cy < 0 ? 0 : (c.cy > R.h - N ? R.h - N : c.cy);
        if (fits(R, N, bx, by)) { best = G.sad(bx, by); ++ie; }
        else { ++oob; }
        int qx = bx * 4, qy = by * 4, q6 = 0;
        if (R.w >= N && R.h >= N) {
            const int scale = search_scale(c.sr, R.w, R.h);
            G.template stage<PAT_LARGE>(scale, 2, bx, by, best, ie, oob);
            G.template stage<PAT_SMALL>(1, 3, bx, by, best, ie, oob);
            G.template stage<PAT_DIAMOND>(1, 4, bx, by, best, ie, oob);
            if (G.lattice_ok(lt, bx, by)) G.qpel(lt, bx, by, qx, qy, best, q6);
            else slow = true;
        } else {
            oob += 6 + 6 + 4;  // one pass of each stage, every candidate off-frame
        }
        if (!slow && G.j == 0) {
            const int i = sh.idx[ci];
            res[i].mv_x = qx;
            res[i].mv_y = qy;
            res[i].sad = best;
            res[i].status = 0;
            if (stats) {
                stats[i].int_evals = ie;
                stats[i].q6_evals = q6;
                stats[i].qbil_evals = 0;
                stats[i].int_oob = oob;
            }
        }
    }
    if (slow && lane % L == 0) {
        // a per-lane slot (no broadcast of the result): safe under divergence
        const int d = atomicAdd(&sh.ndef[di], 1);
        sh.def_ci[di][d] = static_cast<uint8_t>(sh.list[li][first + g]);
        sh.def_x[di][d] = static_cast<int16_t>(bx);
        sh.def_y[di][d] = static_cast<int16_t>(by);
        sh.def_sad[di][d] = best;
        sh.def_ie[di][d] = static_cast<uint8_t>(ie);
        sh.def_oob[di][d] = static_cast<uint8_t>(oob);
    }
}

// Quarter-pel pass of deferred call d (size class di) by one wavefront, with
// its (N+2)^2 anchor window at qw.
template <int N, int WAVES>
__device__ __forceinline__ void run_deferred(TileShared<WAVES> &sh, const PicDev *__restrict__ table,
                                             const PicDev &R, int di, int d, anc_t *qw, int lane,
                                             tcme_result *__restrict__ res,
                                             tcme_stats *__restrict__ stats)
{
    const int ci = sh.def_ci[di][d];
    const tcme_call c = sh.call[ci];
    const PicDev C = table[c.cur_pic];
    Searcher<N, 1> Q(R, C, c.px, c.py, qw, nullptr);
    int32_t best = sh.def_sad[di][d];
    int qx, qy;
    Q.qpel(sh.def_x[di][d], sh.def_y[di][d], qx, qy, best);
    if (lane == 0) {
        const int i = sh.idx[ci];
        res[i].mv_x = qx;
        res[i].mv_y = qy;
        res[i].sad = best;
        res[i].status = 0;
        if (stats) {
            stats[i].int_evals = sh.def_ie[di][d];
            stats[i].q6_evals = Q.cnt.q6;
            stats[i].qbil_evals = Q.cnt.qb;
            stats[i].int_oob = sh.def_oob[di][d];
        }
    }
}

// Quarter-pel pass of deferred 64x64 call d by the whole workgroup (a 64x64
// block held by one wavefront would need ~170 registers; spread over WAVES
// wavefronts each lane holds 64/WAVES samples).  qw: the 66x66 anchor window,
// red: 2 * WAVES * 8 ints for the cross-wavefront sums.  Every wavefront of
// the workgroup calls this with the same d.
template <int WAVES>
__device__ __forceinline__ void run_deferred_wg(TileShared<WAVES> &sh,
                                                const PicDev *__restrict__ table, const PicDev &R,
                                                int d, anc_t *qw, int *red,
                                                tcme_result *__restrict__ res,
                                                tcme_stats *__restrict__ stats)
{
    const int ci = sh.def_ci[0][d];
    const tcme_call c = sh.call[ci];
    const PicDev C = table[c.cur_pic];
    Searcher<64, WAVES> Q(R, C, c.px, c.py, qw, red);
    int32_t best = sh.def_sad[0][d];
    int qx, qy;
    Q.qpel(sh.def_x[0][d], sh.def_y[0][d], qx, qy, best);
    if (Q.tid == 0) {
        const int i = sh.idx[ci];
        res[i].mv_x = qx;
        res[i].mv_y = qy;
        res[i].sad = best;
        res[i].status = 0;
        if (stats) {
            stats[i].int_evals = sh.def_ie[0][d];
            stats[i].q6_evals = Q.cnt.q6;
            stats[i].qbil_evals = Q.cnt.qb;
            stats[i].int_oob = sh.def_oob[0][d];
        }
    }
}

// Every lane adds 1, so a fetch advances the counter by 64 and the item index
// is the first lane's old value / 64.  (A fetch under `if (lane == 0)`
// followed by a broadcast lets the compiler run the other lanes through the
// caller's loop on their own with a stale index; a uniform add compiles to one
// LDS atomic, no branch.)
__device__ __forceinline__ int fetch_item(int *counter)
{
    return __builtin_amdgcn_readfirstlane(atomicAdd(counter, 1)) >> 6;
}

// BIG: the workgroups also take the 32x32 and 64x64 calls (else those are
// left to me_search_kernel, and calls of unsupported sizes to its 64 instance).
template <int WAVES, bool BIG>
__global__ __launch_bounds__(64 * WAVES, BIG && WAVES == 8 ? 4 : 1) void me_tile_kernel(const PicDev *__restrict__ table,
                                                            int npics,
                                                            const tcme_call *__restrict__ calls,
                                                            int n, int chunk,
                                                            tcme_result *__restrict__ res,
                                                            tcme_stats *__restrict__ stats)
{
    __shared__ TileShared<WAVES> sh;
    // chunks from the end of the list first: lists in CTU order end with the
    // picture's partial edge CTUs, whose chunks straddle two CTUs (two runs,
    // twice the staging), and reference-major lists over a most-recent-first
    // picture buffer with the oldest references, whose calls walk furthest
    // and defer most; started last, such chunks set the launch's tail
    const int b = static_cast<int>(gridDim.x) - 1 - tile_slot(blockIdx.x, gridDim.x);
    const int c0 = b * chunk;
    if (c0 >= n) return;
    const int cnt = n - c0 < chunk ? n - c0 : chunk;
    const int lane = threadIdx.x & 63;
    const int wave = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x) >> 6);
    if (wave == 0) {
        // the chunk's valid calls, list order; invalid ones get their status
        int base = 0;
        for (int t0 = 0; t0 < cnt; t0 += 64) {
            const int t = t0 + lane;
            bool live = false;
            tcme_call c{};
            if (t < cnt) {
                c = calls[c0 + t];
                if (BIG || c.blk == 8 || c.blk == 16) {
                    PicDev R{}, C{};
                    const int status = check_call(table, npics, c, R, C);
                    if (status) put_status(res + c0 + t, stats ? stats + c0 + t : nullptr, status);
                    else live = true;
                }
            }
            const uint64_t m = __ballot(live);
            if (live) {
                const int pos = base + __popcll(m & ((1ull << lane) - 1));
                sh.call[pos] = c;
                sh.idx[pos] = c0 + t;
            }
            base += __popcll(m);
        }
        if (lane == 0) sh.nlive = base;
    }
    __syncthreads();
    const int nlive = sh.nlive;
    int pos = 0;
    while (pos < nlive) {
        const int ref = sh.call[pos].ref_pic;
        const int tx = centre_tile(sh.call[pos].cx), ty = centre_tile(sh.call[pos].cy);
        // the run ends at the first later call with another reference or tile
        int end = nlive;
        for (int p0 = pos + 1; p0 < nlive; p0 += 64) {
            const int p = p0 + lane;
            bool brk = false;
            if (p < nlive) {
                brk = sh.call[p].ref_pic != ref || centre_tile(sh.call[p].cx) != tx ||
                      centre_tile(sh.call[p].cy) != ty;
            }
            const uint64_t m = __ballot(brk);
            if (m) { end = p0 + __builtin_ctzll(m); break; }
        }
        const PicDev R = table[ref];
        LumaWin lw;
        lw.w = sh.luma;
        lw.x0 = tx * 64 - kWinMargin < 0 ? 0 : tx * 64 - kWinMargin;
        lw.y0 = ty * 64 - kWinMargin < 0 ? 0 : ty * 64 - kWinMargin;
        lw.x1 = tx * 64 + 64 + kWinMargin > R.w ? R.w : tx * 64 + 64 + kWinMargin;
        lw.y1 = ty * 64 + 64 + kWinMargin > R.h ? R.h : ty * 64 + 64 + kWinMargin;
        if (lw.x1 > lw.x0 && lw.y1 > lw.y0) {
            // rows start 8-byte aligned (x0 is a multiple of 8, pitch of 64);
            // the last dword may run past x1 into the row's padding, the next
            // row or the plane's tail slack
            // 16-byte reads: a row's last one may run up to 15 bytes past x1
            // (row padding, the next row or the plane's tail slack); the LDS
            // row holds kWinPitch / 4 >= 4 * q4 dwords
            const int q4 = (((lw.x1 - lw.x0 + 3) >> 2) + 3) >> 2;
            const int rows = lw.y1 - lw.y0;
            const gptr8 plane = (gptr8)R.y;
            static_assert(4 * ((kWinSide / 4 + 3) / 4) <= kWinPitch / 4, "row of 16-byte reads");
#pragma unroll 4
            for (int i = static_cast<int>(threadIdx.x); i < rows * q4; i += 64 * WAVES) {
                const int r = i / q4, k = i - r * q4;
                const size_t g = static_cast<size_t>(lw.y0 + r) * R.pitch + lw.x0 + 16 * k;
                const u32x4 v = *(gptr128)(plane + g);
                uint32_t *dst = sh.luma + r * (kWinPitch / 4) + 4 * k;
                dst[0] = v.x;
                dst[1] = v.y;
                dst[2] = v.z;
                dst[3] = v.w;
            }
        } else {
            lw.x1 = lw.x0;  // nothing staged: every candidate reads the plane
        }
        LatWin lt;
        lt.w = sh.lat;
        lt.x0 = tx * 64 - kLatMargin < 0 ? 0 : tx * 64 - kLatMargin;
        lt.y0 = ty * 64 - kLatMargin < 0 ? 0 : ty * 64 - kLatMargin;
        lt.x1 = tx * 64 + 64 + kLatMargin > R.w ? R.w : tx * 64 + 64 + kLatMargin;
        lt.y1 = ty * 64 + 64 + kLatMargin > R.h ? R.h : ty * 64 + 64 + kLatMargin;
        if (lt.x1 > lt.x0 && lt.y1 > lt.y0) {
            // two anchors per 16-byte read (x0 and the anchor pitch are even);
            // with an odd width the last read takes one anchor past x1 (the
            // next row or the plane's tail slack) into lattice column x1 - x0
            // < kLatPitch
            const int cols = lt.x1 - lt.x0, rows = lt.y1 - lt.y0;
            const int c2 = (cols + 1) >> 1;
            const gptr128 an2 = (gptr128)R.an;
            static_assert(kLatSide + 1 <= kLatPitch, "odd width's extra column");
#pragma unroll 4
            for (int i = static_cast<int>(threadIdx.x); i < rows * c2; i += 64 * WAVES) {
                const int r = i / c2, c = 2 * (i - r * c2);
                const u32x4 v = an2[(static_cast<size_t>(lt.y0 + r) * R.apitch + lt.x0 + c) >> 1];
                uint32_t *lo = sh.lat + (2 * r) * kLatPitch + c;
                uint32_t *hi = lo + kLatPitch;
                lo[0] = v.x;
                hi[0] = v.y;
                lo[1] = v.z;
                hi[1] = v.w;
            }
        } else {
            lt.x1 = lt.x0;  // empty: every call is deferred
        }
        if (wave == 0) {
            // the run's calls by size class (64, 32, 16, 8), list order
            int nc[4] = {0, 0, 0, 0};
            for (int p0 = pos; p0 < end; p0 += 64) {
                const int p = p0 + lane;
                const int k = p < end ? size_class(sh.call[p].blk) : 4;
                const uint64_t below = (1ull << lane) - 1;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const uint64_t m = __ballot(k == q);
                    if (k == q) sh.list[q][nc[q] + __popcll(m & below)] = static_cast<uint8_t>(p);
                    nc[q] += __popcll(m);
                }
            }
            if (lane == 0) {
#pragma unroll
                for (int q = 0; q < 4; ++q) sh.nlist[q] = nc[q];
                sh.ndef[0] = sh.ndef[1] = sh.ndef[2] = 0;
                sh.next = 0;
            }
        }
        __syncthreads();
        // slots, largest blocks first: 64 (1 call), 32 (4), 16 (8), 8 (8)
        const int n64 = sh.nlist[0], n32 = sh.nlist[1], n16 = sh.nlist[2], n8 = sh.nlist[3];
        constexpr int C64 = Lanes<64>::C, C32 = Lanes<32>::C, C16 = Lanes<16>::C, C8 = Lanes<8>::C;
        const int s64 = (n64 + C64 - 1) / C64, s32 = (n32 + C32 - 1) / C32;
        const int s16 = (n16 + C16 - 1) / C16, s8 = (n8 + C8 - 1) / C8;
        for (;;) {
            int s = fetch_item(&sh.next);
            if constexpr (BIG) {
                if (s < s64) {
                    run_slot<64>(sh, table, R, lw, lt, s, n64 - s, lane, res, stats);
                    continue;
                }
                s -= s64;
                if (s < s32) {
                    run_slot<32>(sh, table, R, lw, lt, C32 * s, n32 - C32 * s, lane, res, stats);
                    continue;
                }
                s -= s32;
            }
            if (s < s16) {
                run_slot<16>(sh, table, R, lw, lt, C16 * s, n16 - C16 * s, lane, res, stats);
                continue;
            }
            s -= s16;
            if (s < s8) {
                run_slot<8>(sh, table, R, lw, lt, C8 * s, n8 - C8 * s, lane, res, stats);
                continue;
            }
            break;
        }
        __syncthreads();
        // Deferred quarter-pel passes; the lattice's LDS now holds anchor
        // windows: first (64+2)^2 for the 64x64 calls, refined one after the
        // other by the whole workgroup; then (16+2)^2 for every wavefront's
        // 8x8/16x16 call and, after those, (32+2)^2 for each of the first few
        // wavefronts' 32x32 call, both kinds refined at the same time.
        const int d64 = sh.ndef[0], d32 = sh.ndef[1], d16 = sh.ndef[2];
        if (d64 + d32 + d16 > 0) {
            constexpr int W32 = 34 * 34, W16 = 18 * 18;
            constexpr int kQw = kLatWords / 2;  // anchors the union holds
            constexpr int P32 = WAVES < kQw / W32 ? WAVES : kQw / W32;
            static_assert(66 * 66 + (2 * WAVES * 8 + 1) / 2 <= kQw, "64x64 window + sums");
            static_assert(P32 >= 1 && WAVES * W16 <= kQw, "deferred windows");
            if constexpr (BIG) {
                // one 64x64 window fits: the whole workgroup refines them in turn
                for (int s = 0; s < d64; ++s) {
                    run_deferred_wg<WAVES>(sh, table, R, s, sh.qw,
                                           reinterpret_cast<int *>(sh.qw + 66 * 66), res, stats);
                    __syncthreads();
                }
                // 32x32 and small calls together: wavefronts below Q32 take the
                // 32x32 ones (windows after every wavefront's small window),
                // then join the others on the small ones
                constexpr int Q32 = (kQw - WAVES * W16) / W32 < P32 ? (kQw - WAVES * W16) / W32 : P32;
                static_assert(Q32 >= 1, "a 32x32 window beside the small ones");
                if (threadIdx.x == 0) { sh.next = 0; sh.next32 = 0; }
                __syncthreads();
                if (wave < Q32) {
                    for (;;) {
                        const int s = fetch_item(&sh.next32);
                        if (s >= d32) break;
                        run_deferred<32>(sh, table, R, 1, s, sh.qw + WAVES * W16 + wave * W32, lane,
                                         res, stats);
                    }
                }
            } else {
                if (threadIdx.x == 0) sh.next = 0;
                __syncthreads();
            }
            for (;;) {
                const int s = fetch_item(&sh.next);
                if (s >= d16) break;
                const int blk = sh.call[sh.def_ci[2][s]].blk;
                if (blk == 16) run_deferred<16>(sh, table, R, 2, s, sh.qw + wave * W16, lane, res, stats);
                else run_deferred<8>(sh, table, R, 2, s, sh.qw + wave * W16, lane, res, stats);
            }
            __syncthreads();
        }
        pos = end;
    }
}

// Anchor plane of a picture (reference src/motion.c:70-107 for the taps):
// {sample, hh, vv, cc} per sample; half samples whose taps would leave the
// frame are written as 0 and never read by the search.  A workgroup of 64 x 4
// threads makes a tile of kAncW x kAncH samples: it stages the tile's luma
// with its apron (2 left / up, 3 right / down; coordinates clamped into the
// frame, and the clamped values never reach a written half sample) in LDS,
// takes the raw horizontal six-tap of every staged row once, and builds the
// diagonal half sample as the vertical six-tap of those (the reference's hv
// order, src/motion.c:94-107).  A thread writes kAncH / 4 rows of one column:
// every row store of a wavefront is 512 contiguous bytes.
constexpr int kAncW = 64, kAncH = 16;
__global__ __launch_bounds__(256) void anchor_kernel(const uint8_t *__restrict__ y, int pitch,
                                                     int w, int h, uint64_t *__restrict__ an,
                                                     int apitch)
{
    constexpr int SR = kAncH + 5, SC = kAncW + 8;  // staged rows / columns
    __shared__ uint8_t lum[SR][SC];                // row y0 - 2 + r, column x0 - 2 + c
    __shared__ int16_t hs[SR][kAncW];              // horizontal six-tap at column x0 + c
    const int x0 = static_cast<int>(blockIdx.x) * kAncW, y0 = static_cast<int>(blockIdx.y) * kAncH;
    const int tx = static_cast<int>(threadIdx.x), ty = static_cast<int>(threadIdx.y);
    const int tid = ty * 64 + tx;
    // every load issued before the first LDS store (a rolled loop waited for
    // each one); slots past the staged area re-read its last sample
    constexpr int NS = (SR * SC + 255) / 256;
    uint8_t v[NS];
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int i = min(tid + 256 * k, SR * SC - 1);
        const int r = i / SC, c = i - r * SC;
        const int gy = min(max(y0 - 2 + r, 0), h - 1), gx = min(max(x0 - 2 + c, 0), w - 1);
        v[k] = y[static_cast<size_t>(gy) * pitch + gx];
    }
#pragma unroll
    for (int k = 0; k < NS; ++k) {
        const int i = tid + 256 * k;
        if (i < SR * SC) lum[i / SC][i % SC] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < (SR * kAncW + 255) / 256; ++k) {
        const int i = tid + 256 * k;
        if (i >= SR * kAncW) break;
        const int r = i / kAncW, c = i - r * kAncW;
        const uint8_t *q = &lum[r][c];  // columns x - 2 .. x + 3
        hs[r][c] = static_cast<int16_t>(tap6(q[0], q[1], q[2], q[3], q[4], q[5]));
    }
    __syncthreads();
    const int x = x0 + tx;
    if (x >= w) return;
    const bool okx = x >= 2 && x + 3 < w;
#pragma unroll
    for (int k = 0; k < kAncH / 4; ++k) {
        const int t = ty + 4 * k;  // tile row; staged rows t .. t + 5 are image rows r - 2 .. r + 3
        const int r = y0 + t;
        if (r >= h) break;
        const bool oky = r >= 2 && r + 3 < h;
        const int hh = okx ? hs[t + 2][tx] : 0;
        int vv = 0, cc = 0;
        if (oky) {
            vv = tap6(lum[t][tx + 2], lum[t + 1][tx + 2], lum[t + 2][tx + 2], lum[t + 3][tx + 2],
                      lum[t + 4][tx + 2], lum[t + 5][tx + 2]);
            if (okx) cc = tap6(hs[t][tx], hs[t + 1][tx], hs[t + 2][tx], hs[t + 3][tx], hs[t + 4][tx],
                               hs[t + 5][tx]);
        }
        const uint32_t lo = static_cast<uint32_t>(lum[t + 2][tx + 2]) | (static_cast<uint32_t>(hh) << 16);
        const uint32_t hi = (static_cast<uint32_t>(vv) & 0xFFFFu) | (static_cast<uint32_t>(cc) << 16);
        an[static_cast<size_t>(r) * apitch + x] = static_cast<uint64_t>(lo) | (static_cast<uint64_t>(hi) << 32);
    }
}

}  // namespace

hipError_t launch_anchors(const PicDev &p, hipStream_t s)
{
    dim3 grid((p.w + kAncW - 1) / kAncW, (p.h + kAncH - 1) / kAncH), block(64, 4);
    hipLaunchKernelGGL(anchor_kernel, grid, block, 0, s, p.y, p.pitch, p.w, p.h,
                       const_cast<uint64_t *>(p.an), p.apitch);
    return hipGetLastError();
}

namespace {

template <int N, int WPC>
hipError_t launch_one(const PicDev *table, int npics, const tcme_call *calls, int n,
                      tcme_result *res, tcme_stats *stats, hipStream_t s)
{
    hipLaunchKernelGGL((me_search_kernel<N, WPC>), dim3(n), dim3(64 * WPC), 0, s, table, npics,
                       calls, n, res, stats);
    return hipGetLastError();
}

}  // namespace

bool valid_waves(int blk, int waves)
{
    switch (blk) {
    case 8: return waves == 1;
    case 16: return waves == 1;
    case 32: return waves == 1 || waves == 2 || waves == 4;
    case 64: return waves == 2 || waves == 4 || waves == 8;
    default: return false;
    }
}

hipError_t launch_tile(int waves, bool big, int chunk, const PicDev *table, int npics,
                       const tcme_call *calls, int n, tcme_result *res, tcme_stats *stats,
                       hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    const int grid = (n + chunk - 1) / chunk;
    if (big && waves == 4)
        hipLaunchKernelGGL((me_tile_kernel<4, true>), dim3(grid), dim3(256), 0, s, table, npics,
                           calls, n, chunk, res, stats);
    else if (big)
        hipLaunchKernelGGL((me_tile_kernel<8, true>), dim3(grid), dim3(512), 0, s, table, npics,
                           calls, n, chunk, res, stats);
    else if (waves == 4)
        hipLaunchKernelGGL((me_tile_kernel<4, false>), dim3(grid), dim3(256), 0, s, table, npics,
                           calls, n, chunk, res, stats);
    else
        hipLaunchKernelGGL((me_tile_kernel<8, false>), dim3(grid), dim3(512), 0, s, table, npics,
                           calls, n, chunk, res, stats);
    return hipGetLastError();
}

hipError_t launch_dropin(const PicDev *table, const tcme_call &c, const uint8_t *blk,
                         tcme_result *res, hipStream_t s)
{
    switch (c.blk) {
    case 8:
        hipLaunchKernelGGL((me_dropin_kernel<8, 1>), dim3(1), dim3(64), 0, s, table, c, blk, res);
        break;
    case 16:
        hipLaunchKernelGGL((me_dropin_kernel<16, 1>), dim3(1), dim3(64), 0, s, table, c, blk, res);
        break;
    case 32:
        hipLaunchKernelGGL((me_dropin_kernel<32, 4>), dim3(1), dim3(256), 0, s, table, c, blk, res);
        break;
    case 64:
        hipLaunchKernelGGL((me_dropin_kernel<64, 8>), dim3(1), dim3(512), 0, s, table, c, blk, res);
        break;
    default:
        if (c.blk < 1) return hipErrorInvalidValue;
        hipLaunchKernelGGL(me_generic_kernel, dim3(1), dim3(256), 0, s, table, c, blk, res);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_list(const PicDev *table, const tcme_call *calls, int n, tcme_result *res,
                       hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(me_list_kernel, dim3(n), dim3(256), 0, s, table, calls, res);
    return hipGetLastError();
}

hipError_t launch_chain(const PicDev *table, const ChainJobHost *jobs, int njobs, int total_ctus,
                        tcme_chain_call *calls, int32_t *counts, hipStream_t s)
{
    if (total_ctus <= 0) return hipSuccess;
    ChainArgs a{};
    a.njobs = njobs;
    for (int i = 0; i < njobs; ++i) {
        ChainJob &j = a.job[i];
        j.cur_pic = jobs[i].cur_pic;
        j.nrefs = jobs[i].nrefs;
        for (int r = 0; r < 4; ++r) j.ref_pic[r] = jobs[i].ref_pic[r];
        j.sr = jobs[i].sr;
        j.lambda = jobs[i].lambda;
        j.first = jobs[i].first;
        j.cols = jobs[i].cols;
    }
    hipLaunchKernelGGL(me_chain_kernel, dim3(total_ctus), dim3(64), 0, s, table, a, calls, counts);
    return hipGetLastError();
}

hipError_t launch_search_n(int blk, int waves, const PicDev *table, int npics,
                           const tcme_call *calls, int n, tcme_result *res, tcme_stats *stats,
                           hipStream_t s)
{
    if (n <= 0) return hipSuccess;
    switch (blk) {
    case 8: return launch_one<8, 1>(table, npics, calls, n, res, stats, s);
    case 16: return launch_one<16, 1>(table, npics, calls, n, res, stats, s);
    case 32:
        if (waves == 1) return launch_one<32, 1>(table, npics, calls, n, res, stats, s);
        if (waves == 4) return launch_one<32, 4>(table, npics, calls, n, res, stats, s);
        return launch_one<32, 2>(table, npics, calls, n, res, stats, s);
    default:
        if (waves == 2) return launch_one<64, 2>(table, npics, calls, n, res, stats, s);
        if (waves == 8) return launch_one<64, 8>(table, npics, calls, n, res, stats, s);
        return launch_one<64, 4>(table, npics, calls, n, res, stats, s);
    }
}

}  // namespace tcme
