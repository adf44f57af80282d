// gfx950 kernels for TCodec's motion compensation (the consumers of each
// search result):
//
//   mc_luma_kernel    tc_inter_predict, reference src/motion.c:493-584, with
//                     get_interp_luma's per-sample ladder (src/motion.c:269-304)
//   mc_chroma_kernel  tc_inter_predict_chroma, src/motion.c:597-632
//
// Every branch of the reference's luma predictor (row copy for an in-frame
// integer MV, whole-block six-tap, per-sample six-tap / bilinear / nearest
// ladder) gives, sample for sample, the ladder's value, so the luma kernel
// computes the ladder directly: six-tap where the sample's 6x6 window is
// inside the frame, bilinear where its 2x2 neighbourhood is, nearest
// (clamped) otherwise.  With the half-sample lattice M[Y][X] (M[2y][2x] the
// sample, M[2y][2x+1] the raw horizontal six-tap hh, M[2y+1][2x] the raw
// vertical one vv, M[2y+1][2x+1] the diagonal cc) and X0 = q >> 1,
// X1 = (q+1) >> 1 per axis at quarter-pel q, the six-tap value is
//   v = clip((avg(M[Y0][X0], M[Y0][X1]) + avg(M[Y1][X0], M[Y1][X1]) + 1) >> 1),
// avg(a,b) = (a+b+1) >> 1 on raw values (src/motion.c:193-260; the h- and
// v-only paths of src/motion.c:133-191 agree after their final clip).  The
// luma kernel computes the lattice values it needs from the luma bytes,
// staged in LDS per block; the chroma kernel reads the plane directly.
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

// Two signed 16-bit lanes of one dword (v_pk_* integer instructions).
typedef short pk16 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ pk16 as_pk16(uint32_t v) { return __builtin_bit_cast(pk16, v); }
__device__ __forceinline__ uint32_t as_u32(pk16 v) { return __builtin_bit_cast(uint32_t, v); }

// ---- luma: aprons of luma bytes staged in LDS ------------------------------

constexpr int kMcChunk = 128;     // descriptors per workgroup at most
constexpr int kMcWaveLds = 3328;  // bytes of staged aprons per wavefront (>= every slot's)

typedef unsigned int u32x4a4 __attribute__((ext_vector_type(4), aligned(4)));
typedef const __attribute__((address_space(1))) u32x4a4 *gptr128u;
typedef unsigned int u32x4l __attribute__((ext_vector_type(4)));

// A slot: G blocks of one size N (64x64 blocks: a band of 16 rows each),
// L lanes each, one task (2 columns x 8 rows) per lane.  The apron of a
// band, rows r0-2 .. r0+BR+2 of the block's reference region and columns
// from (fx-2) rounded down to a multiple of 4, takes ROWS rows of P bytes
// (N+5 samples plus up to 3 of alignment, in 16-byte segments).
template <int N>
struct McShape {
    static constexpr int BR = N == 64 ? 16 : N;  // block rows per slot
    static constexpr int P = 16 * ((N + 8 + 15) / 16);
    static constexpr int P16 = P / 16;
    static constexpr int ROWS = BR + 5;
    static constexpr int PAIRS = N / 2, BANDS = BR / 8;
    static constexpr int L = PAIRS * BANDS;      // lanes per block band
    static constexpr int G = 64 / L;             // block bands per slot
    static constexpr int S = ROWS * P16;         // segments per apron
    static constexpr int SPL = (S + L - 1) / L;  // segments per lane
    static_assert(L <= 64 && 64 % L == 0, "lanes");
    static_assert(G * ROWS * P <= kMcWaveLds, "slot does not fit");
};

// Per-descriptor facts a slot needs, in LDS (48 bytes).
struct McInfo {
    uint64_t src;     // the plane's first byte
    uint64_t dst;     // the block's first output byte
    int32_t pitch;    // plane pitch
    int32_t frac;     // dx | dy << 2 | ((fx - 2) & 3) << 4 | edge << 6
    int32_t ostride;  // output row stride
    int32_t pic;      // picture id
    int32_t fx, fy;   // integer part of the MV
    int32_t w, h;     // picture size
};

// The raw six-tap (1, -5, 20, 20, -5, 1) of six 16-bit values, (+16) >> 5,
// two at a time (the sums stay within [-6710, 14886]: no lane wraps).
__device__ __forceinline__ pk16 v6(pk16 a0, pk16 a1, pk16 a2, pk16 a3, pk16 a4, pk16 a5)
{
    const pk16 c16 = {16, 16}, c5 = {5, 5}, c20 = {20, 20};
    return ((a0 + a5) - c5 * (a1 + a4) + c20 * (a2 + a3) + c16) >> 5;
}

// get_interp_luma's ladder at one sample with every load issued up front:
// anchors and samples at the four clamped corners and the nearest sample
// (each branch reads only coordinates that are in the frame, where the
// clamped ones equal its own).
__device__ __forceinline__ int luma_px_flat(const PicDev &R, int x, int y, int dx, int dy)
{
    const int cx0 = min(max(x, 0), R.w - 1), cx1 = min(max(x + 1, 0), R.w - 1);
    const int cy0 = min(max(y, 0), R.h - 1), cy1 = min(max(y + 1, 0), R.h - 1);
    const int nx = min(max(x + (dx >= 2 ? 1 : 0), 0), R.w - 1);
    const int ny = min(max(y + (dy >= 2 ? 1 : 0), 0), R.h - 1);
    const gptr8 Y = (gptr8)R.y;
    const gptr64 A = (gptr64)R.an;
    const size_t r0 = static_cast<size_t>(cy0), r1 = static_cast<size_t>(cy1);
    const uint64_t a00 = A[r0 * R.apitch + cx0], a10 = A[r0 * R.apitch + cx1];
    const uint64_t a01 = A[r1 * R.apitch + cx0], a11 = A[r1 * R.apitch + cx1];
    const int p00 = Y[r0 * R.pitch + cx0], p10 = Y[r0 * R.pitch + cx1];
    const int p01 = Y[r1 * R.pitch + cx0], p11 = Y[r1 * R.pitch + cx1];
    const int pn = Y[static_cast<size_t>(ny) * R.pitch + nx];
    if (x >= 2 && x + 3 < R.w && y >= 2 && y + 3 < R.h)
        return (dx | dy) == 0 ? p00 : lattice_px(a00, a10, a01, a11, dx, dy);
    if (x >= 0 && y >= 0 && x + 1 < R.w && y + 1 < R.h)
        return lerp4(lerp4(p00, p10, dx), lerp4(p01, p11, dx), dy);
    return pn;
}

// What a block asks of the task code: KIND bit 0, dx even (X0 = X1: one
// lattice column per row); bit 1, dy = 0 (no vertical half samples, apron
// rows 2..9 only).  Kind 0 is the general case and serves any mix.
__device__ __forceinline__ int mc_kind(int dx, int dy)
{
    return ((dx & 1) == 0 ? 1 : 0) | (dy == 0 ? 2 : 0);
}

// Rows r0 .. r0+7 of columns c, c+1 of a staged block (apron at lds, row
// pitch P, apron column 0 at byte a of each row) at fraction (dx, dy), the
// two columns in the 16-bit halves of res[r].  Each of the task's 13 apron
// rows gives s(x), s(x+1), s(x+2) and the raw hh(x), hh(x+1) (signed-byte
// dot products: bytes b-128 against the taps, 128*32 added back); the row's
// lattice values at X0 and X1 (s or hh, by dx) are e0 and e1.  The odd
// lattice row between rows y and y+1 holds the vertical six-taps of those
// columns (vv of a sample column, cc of an hh column), so with hE(y) =
// avg(e0, e1) on row y and hO(y) = avg(v6(e0), v6(e1)) the value is
// clip(avg(top, bot)) with top = hE(y) or hO(y) and bot = hE(y), hO(y) or
// hE(y+1), by dy.  KIND (mc_kind) drops what the block does not need.
template <int KIND>
__device__ __forceinline__ void mc_task(const uint8_t *lds, int P, int a, int dx, int dy, int c,
                                        int r0, uint32_t (&res)[8])
{
    constexpr bool ONE = (KIND & 1) != 0, FLAT = (KIND & 2) != 0;
    constexpr int RB = FLAT ? 2 : 0, RE = FLAT ? 10 : 13;  // apron rows used
    const int off = a + c, sh = off & 3;
    const uint8_t *b = lds + r0 * P + (off & ~3);
    constexpr uint32_t kM = 0x80808080u;
    const int X0 = dx >> 1, X1 = (dx + 1) >> 1;
    const int K0 = static_cast<int>(0x1414FB01u), K1 = 0x000001FB;
    const bool xa = X0 != 0, xb0 = X1 == 0, xb2 = X1 == 2;
    uint32_t E0[13], E1[13];
#pragma unroll
    for (int i = RB; i < RE; ++i) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(b + i * P);
        const uint32_t d0 = w[0], d1 = w[1], d2 = w[2];
        const uint32_t w0 = __builtin_amdgcn_alignbyte(d1, d0, sh);
        const uint32_t w1 = __builtin_amdgcn_alignbyte(d2, d1, sh);
        const uint32_t u0 = w0 ^ kM, u1 = w1 ^ kM;
        const int h0 = __builtin_amdgcn_sdot4(static_cast<int>(u0), K0,
                                              __builtin_amdgcn_sdot4(static_cast<int>(u1), K1,
                                                                     4112, false),
                                              false);
        const int h1 = __builtin_amdgcn_sdot4(
            static_cast<int>(__builtin_amdgcn_alignbyte(u1, u0, 1)), K0,
            __builtin_amdgcn_sdot4(static_cast<int>(u1 >> 8), K1, 4112, false), false);
        // (column c, column c+1) in the two 16-bit halves
        const uint32_t sp = __builtin_amdgcn_perm(w1, w0, 0x0c030c02u);  // s(x), s(x+1)
        const uint32_t hp = as_u32(as_pk16(__builtin_amdgcn_perm(static_cast<uint32_t>(h1),
                                                                 static_cast<uint32_t>(h0),
                                                                 0x05040100u)) >> 5);
        E0[i] = xa ? hp : sp;
        if constexpr (!ONE) {
            const uint32_t sq = __builtin_amdgcn_perm(w1, w0, 0x0c040c03u);  // s(x+1), s(x+2)
            E1[i] = xb0 ? sp : (xb2 ? sq : hp);
        }
    }
    const pk16 one = {1, 1}, lo = {0, 0}, hi = {255, 255};
    if constexpr (FLAT) {
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            pk16 v = as_pk16(E0[r + 2]);
            if constexpr (!ONE) v = (v + as_pk16(E1[r + 2]) + one) >> 1;
            res[r] = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi));
        }
    } else {
        const bool yt = dy >= 2, y1e = dy == 0, y1f = dy == 3;
        pk16 hE = as_pk16(E0[2]);
        if constexpr (!ONE) hE = (hE + as_pk16(E1[2]) + one) >> 1;
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            pk16 hO = v6(as_pk16(E0[r]), as_pk16(E0[r + 1]), as_pk16(E0[r + 2]),
                         as_pk16(E0[r + 3]), as_pk16(E0[r + 4]), as_pk16(E0[r + 5]));
            pk16 hF = as_pk16(E0[r + 3]);
            if constexpr (!ONE) {
                const pk16 o1 = v6(as_pk16(E1[r]), as_pk16(E1[r + 1]), as_pk16(E1[r + 2]),
                                   as_pk16(E1[r + 3]), as_pk16(E1[r + 4]), as_pk16(E1[r + 5]));
                hO = (hO + o1 + one) >> 1;
                hF = (hF + as_pk16(E1[r + 3]) + one) >> 1;
            }
            const pk16 top = yt ? hO : hE;
            const pk16 bot = y1e ? hE : (y1f ? hF : hO);
            const pk16 v = (top + bot + one) >> 1;
            res[r] = as_u32(__builtin_elementwise_min(__builtin_elementwise_max(v, lo), hi));
            hE = hF;
        }
    }
}

// A lane's task of a slot: 2 columns x 8 rows of one block, its apron staged
// at lds (row pitch P).
struct McTask {
    const uint8_t *lds;
    int P, rows;   // LDS row pitch, staged rows
    int rb, c, r0; // band's first block row, the task's column pair and rows
    bool act;      // the lane's group holds a block
};

// Stages one slot: up to G blocks of size N (descriptor numbers idx[0..cnt)
// of the workgroup's chunk; for N = 64 the band of 16 rows `band`), L lanes
// each.  The lanes stage their block's apron in the wavefront's own LDS
// region and return their task; `edge` tells whether a block of the slot
// has an apron reaching past the picture.
//   no such block: 16-byte loads;
//   else: dword loads with the row clamped into the picture and the dword
//         into the row, its bytes permuted so that the copy's byte at
//         column X is the plane's at clamp(X, 0, w - 1).
template <int N>
__device__ __forceinline__ McTask mc_stage(const McInfo *info, const uint8_t *idx, int cnt,
                                           int band, uint8_t *lds, int lane, McInfo &m, bool &edge)
{
    using S = McShape<N>;
    const int g = lane / S::L, li = lane % S::L;
    const bool act = g < cnt;
    m = info[idx[act ? g : 0]];
    edge = __ballot(act && (m.frac & 64)) != 0;
    uint8_t *blds = lds + g * S::ROWS * S::P;
    const int rb = band * S::BR;  // the band's first block row
    const int xs = (m.fx - 2) & ~3;
    if (!edge) {
        const gptr8 src = (gptr8)m.src + static_cast<size_t>(m.fy - 2 + rb) * m.pitch + xs;
        // every lane loads (segments past the apron re-read its last one,
        // the groups without a block read the slot's first block), so the
        // loads are issued together
        u32x4l v[S::SPL];
#pragma unroll
        for (int i = 0; i < S::SPL; ++i) {
            const int s = min(li + S::L * i, S::S - 1);
            const int row = s / S::P16, q = s % S::P16;
            const u32x4a4 t = *(gptr128u)(src + static_cast<size_t>(row) * m.pitch + 16 * q);
            v[i] = u32x4l{t.x, t.y, t.z, t.w};
        }
#pragma unroll
        for (int i = 0; i < S::SPL; ++i) {
            const int s = min(li + S::L * i, S::S - 1);
            const int row = s / S::P16, q = s % S::P16;
            *reinterpret_cast<u32x4l *>(blds + row * S::P + 16 * q) = v[i];
        }
    } else {
        constexpr int D4 = S::P / 4, DS = S::ROWS * D4, DPL = (DS + S::L - 1) / S::L;
        const gptr8 src = (gptr8)m.src;
        uint32_t v[DPL];
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const int s = min(li + S::L * i, DS - 1);
            const int row = s / D4, dc = s % D4;
            const int yy = min(max(m.fy - 2 + rb + row, 0), m.h - 1);
            const int xx = min(max(xs + 4 * dc, 0), max(m.w - 4, 0));
            v[i] = *(const __attribute__((address_space(1))) uint32_t *)(
                src + static_cast<size_t>(yy) * m.pitch + xx);
        }
#pragma unroll
        for (int i = 0; i < DPL; ++i) {
            const int s = min(li + S::L * i, DS - 1);
            const int row = s / D4, dc = s % D4;
            // byte j of the copy is the plane's byte at column
            // clamp(x0 + j, 0, w - 1), which lies in the loaded dword
            const int x0 = xs + 4 * dc;
            const int xx = min(max(x0, 0), max(m.w - 4, 0));
            uint32_t sel = 0;
#pragma unroll
            for (int j = 0; j < 4; ++j)
                sel |= static_cast<uint32_t>(min(max(x0 + j, 0), m.w - 1) - xx) << (8 * j);
            *reinterpret_cast<uint32_t *>(blds + row * S::P + 4 * dc) =
                __builtin_amdgcn_perm(0u, v[i], sel);
        }
    }
    McTask t;
    t.lds = blds;
    t.P = S::P;
    t.rows = S::ROWS;
    t.rb = rb;
    t.c = 2 * (li % S::PAIRS);
    t.r0 = 8 * (li / S::PAIRS);
    t.act = act;
    return t;
}

// Computes a staged slot's tasks and stores them (a row's two samples in
// one 16-bit store when aligned).  The slot's blocks share one code path:
// the specialised one of their kind when they all have the same (mc_kind),
// else the general one.  `edge` (the slot's blocks may reach past the
// picture, staged with clamped coordinates): a sample whose 6x6 window is
// in the picture keeps the six-tap value, any other one takes the ladder's
// bilinear (2x2 in the picture) or nearest value, both read from the copy.
__device__ __forceinline__ void mc_compute(const McTask &t, const McInfo &m, bool edge)
{
    // the wavefront's LDS writes before its reads (one wavefront owns the region)
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (t.act) {
        const int dx = m.frac & 3, dy = (m.frac >> 2) & 3, a = (m.frac >> 4) & 3;
        const int kind = mc_kind(dx, dy);
        const int k0 = __builtin_amdgcn_readfirstlane(kind);
        const int kk = __ballot(kind != k0) ? 0 : k0;
        uint32_t res[8];
        switch (kk) {
        case 1: mc_task<1>(t.lds, t.P, a, dx, dy, t.c, t.r0, res); break;
        case 2: mc_task<2>(t.lds, t.P, a, dx, dy, t.c, t.r0, res); break;
        case 3: mc_task<3>(t.lds, t.P, a, dx, dy, t.c, t.r0, res); break;
        default: mc_task<0>(t.lds, t.P, a, dx, dy, t.c, t.r0, res); break;
        }
        const gout8 oc = (gout8)m.dst + static_cast<int64_t>(t.rb + t.r0) * m.ostride + t.c;
        if (edge && (m.frac & 64)) {
            // samples whose 6x6 window leaves the picture (bit 2r + column)
            // take the ladder's bilinear value (2x2 in the picture) or its
            // nearest sample, both from the copy, which holds the plane with
            // clamped coordinates; patched into the packed results
            const int x = m.fx + t.c, y = m.fy + t.rb + t.r0;
            uint32_t fix = 0, bil = 0;
#pragma unroll
            for (int r = 0; r < 8; ++r) {
#pragma unroll
                for (int cc = 0; cc < 2; ++cc) {
                    const int xk = x + cc, yk = y + r;
                    if (!(xk >= 2 && xk + 3 < m.w && yk >= 2 && yk + 3 < m.h))
                        fix |= 1u << (2 * r + cc);
                    if (xk >= 0 && xk + 1 < m.w && yk >= 0 && yk + 1 < m.h)
                        bil |= 1u << (2 * r + cc);
                }
            }
            if (fix) {
                // rows y .. y+8, columns x .. x+2 of the copy, read together
                const uint8_t *q = t.lds + (t.r0 + 2) * t.P + t.c + a + 2;
                int p[9][3];
#pragma unroll
                for (int i = 0; i < 9; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) p[i][j] = q[i * t.P + j];
                const int ox = dx >= 2 ? 1 : 0, oy = dy >= 2 ? 1 : 0;
#pragma unroll
                for (int k = 0; k < 16; ++k) {
                    const int r = k >> 1, cc = k & 1;
                    const int b = lerp4(lerp4(p[r][cc], p[r][cc + 1], dx),
                                        lerp4(p[r + 1][cc], p[r + 1][cc + 1], dx), dy);
                    const int nr = oy ? p[r + 1][cc] : p[r][cc];
                    const int nn = oy ? p[r + 1][cc + 1] : p[r][cc + 1];
                    const int v = ((bil >> k) & 1) ? b : (ox ? nn : nr);
                    if ((fix >> k) & 1)
                        res[r] = cc ? (res[r] & 0xffffu) | (static_cast<uint32_t>(v) << 16)
                                    : (res[r] & 0xffff0000u) | static_cast<uint32_t>(v);
                }
            }
        }
        if (((m.dst + t.c) & 1) == 0 && (m.ostride & 1) == 0) {
            // both columns of a row in one 16-bit store
            typedef __attribute__((address_space(1))) uint16_t *gout16;
#pragma unroll
            for (int r = 0; r < 8; ++r)
                *(gout16)(oc + static_cast<int64_t>(r) * m.ostride) =
                    static_cast<uint16_t>((res[r] & 0xffu) | ((res[r] >> 8) & 0xff00u));
        } else {
#pragma unroll
            for (int r = 0; r < 8; ++r) {
                oc[static_cast<int64_t>(r) * m.ostride] = static_cast<uint8_t>(res[r]);
                oc[static_cast<int64_t>(r) * m.ostride + 1] = static_cast<uint8_t>(res[r] >> 16);
            }
        }
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Workgroup b (4 wavefronts) takes up to kMcChunk consecutive descriptors (the
// chunks of one XCD are contiguous: workgroup b runs on XCD b % 8, so the
// blocks of one CTU, whose quadtree levels read the same region, meet in
// one L2).  It validates them and sorts the blocks of sizes 8, 16, 32 and
// 64 into slots of one size (64/L blocks, L lanes each), each size's list
// ordered by whether the block's apron reaches past the picture and by
// kind (mc_kind), so that a slot's blocks mostly share their staging and
// code path; blocks of other sizes go into slots of 128 samples.  The
// wavefronts take slots from a counter in LDS, the slowest first.  A slot
// is one round trip to memory: its aprons come into the wavefront's LDS and
// every half sample is computed from them.
__global__ __launch_bounds__(256) void mc_luma_kernel(const PicDev *__restrict__ table, int npics,
                                                      const tcme_pred *__restrict__ preds, int n,
                                                      uint8_t *__restrict__ out, int64_t out_bytes,
                                                      int32_t *__restrict__ status)
{
    // classes: 0-3 sizes 64/32/16/8, 4 other sizes; keys: inner << 2 | kind
    __shared__ McInfo info[kMcChunk];
    __shared__ uint8_t lists[5][kMcChunk];
    __shared__ int counts[5];
    __shared__ int kcount[5][8];
    __shared__ int next;
    __shared__ int oslot[kMcChunk + 1];
    __shared__ u32x4l stage[4 * kMcWaveLds / 16 + 1];
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int kXcd = 8;
    const int g = static_cast<int>(gridDim.x), bid = static_cast<int>(blockIdx.x);
    const int q = g / kXcd, rr = g % kXcd, xc = bid % kXcd;
    const int chunk = xc * q + (xc < rr ? xc : rr) + bid / kXcd;   // XCD xc's chunks are contiguous
    // an even split: at most kMcChunk descriptors a workgroup, and at least
    // 8 workgroups a CU, so that a short list of large blocks still spreads
    const int64_t c0 = static_cast<int64_t>(n) * chunk / g;
    const int cnt = static_cast<int>(static_cast<int64_t>(n) * (chunk + 1) / g - c0);
    if (tid < 40) kcount[tid >> 3][tid & 7] = 0;
    if (tid == 0) next = 0;
    __syncthreads();
    int cl = -1, key = 0, rank = 0;
    if (tid < cnt) {
        const tcme_pred p = preds[c0 + tid];
        PicDev R{};
        int st = 0;
        if (p.pic < 0 || p.pic >= npics) {
            st = TCME_E_PIC;
        } else {
            R = table[p.pic];
            if (!R.y || !R.an) st = TCME_E_PIC;
            else if (p.blk < 1 || p.blk > kMaxBlk) st = TCME_E_BLK;
            else if (!span_fits(p.out_off, p.blk, p.blk, p.out_stride, out_bytes)) st = TCME_E_ARG;
        }
        if (status) status[c0 + tid] = st;
        if (st == 0) {
            const int fx = p.mv_x >> 2, fy = p.mv_y >> 2, nb = p.blk;
            const bool inner = fx >= 2 && fy >= 2 && fx + nb + 3 <= R.w && fy + nb + 3 <= R.h;
            const int sz = nb == 64 ? 0 : nb == 32 ? 1 : nb == 16 ? 2 : nb == 8 ? 3 : 4;
            McInfo m;
            m.src = reinterpret_cast<uint64_t>(R.y);
            m.dst = reinterpret_cast<uint64_t>(out + p.out_off);
            m.pitch = R.pitch;
            m.frac = (p.mv_x & 3) | ((p.mv_y & 3) << 2) | (((fx - 2) & 3) << 4) |
                     (inner ? 0 : 64);
            m.ostride = p.out_stride;
            m.pic = p.pic;
            m.fx = fx;
            m.fy = fy;
            m.w = R.w;
            m.h = R.h;
            info[tid] = m;
            cl = sz;
            key = (inner ? 4 : 0) | mc_kind(p.mv_x & 3, p.mv_y & 3);
            rank = atomicAdd(&kcount[sz][key], 1);
        }
    }
    __syncthreads();
    if (cl >= 0) {
        int pos = rank;
        for (int k = 0; k < key; ++k) pos += kcount[cl][k];
        lists[cl][pos] = static_cast<uint8_t>(tid);
    }
    if (tid < 5) {
        int c = 0;
        for (int k = 0; k < 8; ++k) c += kcount[tid][k];
        counts[tid] = c;
    }
    __syncthreads();
    const int no = counts[4];
    if (no && wave == 0) {
        // other-size blocks' first slots (128 samples a slot): a prefix sum
        // over the list, 64 entries at a time
        int base = 0;
        for (int k0 = 0; k0 < no; k0 += 64) {
            const int k = k0 + lane;
            int v = 0;
            if (k < no) {
                const int nb = preds[c0 + lists[4][k]].blk;
                v = (nb * nb + 127) >> 7;
            }
            int incl = v;
#pragma unroll
            for (int d = 1; d < 64; d <<= 1) {
                const int t = __builtin_amdgcn_ds_bpermute((lane >= d ? lane - d : lane) << 2, incl);
                if (lane >= d) incl += t;
            }
            if (k < no) oslot[k] = base + incl - v;
            base += __builtin_amdgcn_readlane(incl, 63);
        }
        if (lane == 0) oslot[no] = base;
    }
    __syncthreads();
    constexpr int G32 = McShape<32>::G, G16 = McShape<16>::G, G8 = McShape<8>::G;
    const int n64 = counts[0], n32 = counts[1], n16 = counts[2], n8 = counts[3];
    // slots in the order they are taken: other sizes, 64 (4 bands a block),
    // 32, 16, 8; within a size the blocks that reach past the picture, which
    // cost the most, lead the list
    const int so = no ? oslot[no] : 0, s64 = 4 * n64, s32 = (n32 + G32 - 1) / G32,
              s16 = (n16 + G16 - 1) / G16, s8 = (n8 + G8 - 1) / G8;
    const int slots = so + s64 + s32 + s16 + s8;
    uint8_t *lds = reinterpret_cast<uint8_t *>(stage) + wave * kMcWaveLds;
    for (;;) {
        int s = __builtin_amdgcn_readfirstlane(atomicAdd(&next, 1)) >> 6;
        if (s >= slots) break;
        // the lane index passes through an empty asm each round, so the
        // compiler cannot hoist the slot shapes' lane-derived offsets out of
        // the loop and hold them all live
        int ln = lane;
        asm volatile("" : "+v"(ln));
        if (s < so) {
            // 128 samples of a block of another size, two per lane
            int k = 0;  // the last block whose first slot is <= s
            for (int sft = kMcChunk / 2; sft; sft >>= 1)
                if (k + sft < no && oslot[k + sft] <= s) k += sft;
            const McInfo m = info[lists[4][k]];
            const PicDev R = table[m.pic];
            const int dx = m.frac & 3, dy = (m.frac >> 2) & 3;
            const int nb = preds[c0 + lists[4][k]].blk, tot = nb * nb;
            const gout8 o = (gout8)m.dst;
            const int u0 = (s - oslot[k]) * 128;
            const int ua = u0 + ln, ub = u0 + 64 + ln;
            const int ca = ua % nb, ra = min(ua / nb, nb - 1);
            const int cb = ub % nb, rb = min(ub / nb, nb - 1);
            const int va = luma_px_flat(R, m.fx + ca, m.fy + ra, dx, dy);
            const int vb = luma_px_flat(R, m.fx + cb, m.fy + rb, dx, dy);
            if (ua < tot) o[static_cast<int64_t>(ra) * m.ostride + ca] = static_cast<uint8_t>(va);
            if (ub < tot) o[static_cast<int64_t>(rb) * m.ostride + cb] = static_cast<uint8_t>(vb);
            continue;
        }
        s -= so;
        McInfo m;
        McTask t;
        bool edge;
        if (s < s64) {
            t = mc_stage<64>(info, lists[0] + (s >> 2), 1, s & 3, lds, ln, m, edge);
        } else if ((s -= s64) < s32) {
            t = mc_stage<32>(info, lists[1] + G32 * s, min(G32, n32 - G32 * s), 0, lds, ln, m, edge);
        } else if ((s -= s32) < s16) {
            t = mc_stage<16>(info, lists[2] + G16 * s, min(G16, n16 - G16 * s), 0, lds, ln, m, edge);
        } else {
            s -= s16;
            t = mc_stage<8>(info, lists[3] + G8 * s, min(G8, n8 - G8 * s), 0, lds, ln, m, edge);
        }
        mc_compute(t, m, edge);
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

// Rows [r0, r0 + 4) x columns [c, c + 4) of a chroma block, plane y / pitch
// (w x h), at 1/8-sample (mvx, mvy).  Each of the 5 source rows is one
// 8-byte load from a dword boundary (the plane's rows start 64-byte aligned
// and end in padding or tail slack).  Without clamping (every sample's 2x2
// neighbourhood inside the plane) the load starts at the dword below column
// x and a 64-bit shift brings column x to byte 0; with clamping the row is
// clamped into the plane, the load starts at the dword below clamp(x), and
// a byte permute picks the bytes at clamp(x + j, 0, w - 1), all within the
// 8 loaded.  A sample is two unsigned byte dot products (v_dot4_u32_u8) of
// the rows' bytes against the weights (8-fx)(8-fy), fx(8-fy) and (8-fx)fy,
// fx fy placed at its columns, the rounding 32 carried in, and a row's four
// samples go out as one dword when the output is 4-byte aligned.
__device__ __forceinline__ void chroma_tile(const uint8_t *y, int pitch, int w, int h, bool clamp,
                                            int mvx, int mvy, int c, int r0, uint8_t *o, int os)
{
    typedef unsigned int u32x2a4 __attribute__((ext_vector_type(2), aligned(4)));
    typedef const __attribute__((address_space(1))) u32x2a4 *gptr64u;
    const int xi = mvx >> 3, yi = mvy >> 3, fx = mvx & 7, fy = mvy & 7;
    const int x = xi + c;
    uint32_t lo[5], l1[5];  // bytes x .. x+3 and x+1 .. x+4 of each row
    if (!clamp) {
        const int sh = x & 3;
        const gptr8 rp = (gptr8)y + static_cast<size_t>(yi + r0) * pitch + (x & ~3);
        uint64_t raw[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const u32x2a4 d = *(gptr64u)(rp + static_cast<size_t>(r) * pitch);
            raw[r] = static_cast<uint64_t>(d.x) | (static_cast<uint64_t>(d.y) << 32);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const uint64_t t = raw[r] >> (8 * sh);
            lo[r] = static_cast<uint32_t>(t);
            l1[r] = static_cast<uint32_t>(t >> 8);
        }
    } else {
        const int base = min(max(x, 0), w - 1) & ~3;
        uint32_t sel = 0, sel1 = 0;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            const uint32_t sj = static_cast<uint32_t>(min(max(x + j, 0), w - 1) - base);  // 0..7
            if (j < 4) sel |= sj << (8 * j);
            if (j > 0) sel1 |= sj << (8 * (j - 1));
        }
        u32x2a4 d[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const int yy = min(max(yi + r0 + r, 0), h - 1);
            d[r] = *(gptr64u)((gptr8)y + static_cast<size_t>(yy) * pitch + base);
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            lo[r] = __builtin_amdgcn_perm(d[r].y, d[r].x, sel);
            l1[r] = __builtin_amdgcn_perm(d[r].y, d[r].x, sel1);
        }
    }
    const uint32_t wa = static_cast<uint32_t>((8 - fx) * (8 - fy)) | (static_cast<uint32_t>(fx * (8 - fy)) << 8);
    const uint32_t wc = static_cast<uint32_t>((8 - fx) * fy) | (static_cast<uint32_t>(fx * fy) << 8);
    const gout8 oc = (gout8)o + static_cast<int64_t>(r0) * os + c;
    const bool al = ((reinterpret_cast<uint64_t>(o) + c) & 3) == 0 && (os & 3) == 0;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t v0 = __builtin_amdgcn_udot4(lo[r], wa, __builtin_amdgcn_udot4(lo[r + 1], wc, 32u, false), false) >> 6;
        const uint32_t v1 = __builtin_amdgcn_udot4(lo[r], wa << 8, __builtin_amdgcn_udot4(lo[r + 1], wc << 8, 32u, false), false) >> 6;
        const uint32_t v2 = __builtin_amdgcn_udot4(lo[r], wa << 16, __builtin_amdgcn_udot4(lo[r + 1], wc << 16, 32u, false), false) >> 6;
        const uint32_t v3 = __builtin_amdgcn_udot4(l1[r], wa << 16, __builtin_amdgcn_udot4(l1[r + 1], wc << 16, 32u, false), false) >> 6;
        const gout8 orow = oc + static_cast<int64_t>(r) * os;
        if (al) {
            *(__attribute__((address_space(1))) uint32_t *)orow = v0 | (v1 << 8) | (v2 << 16) | (v3 << 24);
        } else {
            orow[0] = static_cast<uint8_t>(v0);
            orow[1] = static_cast<uint8_t>(v1);
            orow[2] = static_cast<uint8_t>(v2);
            orow[3] = static_cast<uint8_t>(v3);
        }
    }
}

// Chroma at 1/8 sample (tc_inter_predict_chroma, src/motion.c:597-632), the
// same work split as the luma kernel: a chunk of the list per workgroup,
// taken 256 descriptors at a time (64 a wavefront); blocks of 4x4 .. 32x32
// (with or without edge clamping) are cut into tasks of 4 x 4 samples
// (m * m / 16 per block), numbered in a row over the 256 descriptors (a
// prefix sum over each wavefront's lanes, then over the wavefronts) and
// taken by the workgroup's 256 lanes 256 at a time, each finding its block
// by a binary search over the prefix sums in LDS (chroma_tile: one round trip
// to memory per pass); blocks of other sizes are predicted one per wavefront,
// each wavefront taking those of its own 64 descriptors.  (Before: tasks of one
// column x 4 rows from byte loads, numbered over 64 descriptors, 26 us for a
// 1080p frame's 343,200 blocks; 4 x 4 tasks numbered over 64 descriptors,
// 21.7 us, a run of 4x4 blocks then filling one wavefront of the four.
// Quads of consecutive blocks per wavefront, four 4x4 blocks at 16 lanes
// each and larger ones one after another, took 63 us, most of it in the
// 16x16 and 32x32 blocks' rounds of loads.)
struct ChromaTask {
    uint64_t y, o;   // plane, block's output
    int mvx, mvy;
    int pitch, os;
    int w, h;        // plane size
    int kl;          // log2 of the block size
    int clamp;       // a sample's 2x2 neighbourhood may leave the plane
};

__global__ __launch_bounds__(256) void mc_chroma_kernel(const PicDev *__restrict__ table,
                                                        int npics,
                                                        const tcme_pred *__restrict__ preds,
                                                        int n, uint8_t *__restrict__ out,
                                                        int64_t out_bytes,
                                                        int32_t *__restrict__ status)
{
    __shared__ ChromaTask info[256];
    __shared__ int excl_s[256];      // first task of each descriptor
    __shared__ int wtot[4];          // tasks per wavefront
    const int tid = static_cast<int>(threadIdx.x), lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    constexpr int kXcd = 8;
    const int g = static_cast<int>(gridDim.x), bid = static_cast<int>(blockIdx.x);
    const int q = g / kXcd, rr = g % kXcd, x = bid % kXcd;
    const int chunk = x * q + (x < rr ? x : rr) + bid / kXcd;   // XCD x's chunks are contiguous
    const int64_t c0 = static_cast<int64_t>(n) * chunk / g;
    const int64_t c1 = static_cast<int64_t>(n) * (chunk + 1) / g;
    PicDev R{};
    int cached = -1;
    for (int64_t w0 = c0; w0 < c1; w0 += 256) {
        const int64_t i = w0 + tid;  // this lane's descriptor
        tcme_pred p{};
        bool ok = false;
        ChromaTask ct{};
        if (i < c1) {
            p = preds[i];
            const int lg = p.blk == 4 ? 2 : p.blk == 8 ? 3 : p.blk == 16 ? 4 : p.blk == 32 ? 5 : 0;
            if (lg && p.pic >= 0 && p.pic < npics) {
                const PicDev Rk = table[p.pic];
                const int xi = p.mv_x >> 3, yi = p.mv_y >> 3;
                ok = Rk.y && span_fits(p.out_off, p.blk, p.blk, p.out_stride, out_bytes);
                ct.clamp = !(xi >= 0 && yi >= 0 && xi + p.blk + 1 <= Rk.w && yi + p.blk + 1 <= Rk.h);
                ct.w = Rk.w;
                ct.h = Rk.h;
                ct.y = reinterpret_cast<uint64_t>(Rk.y);
                ct.o = reinterpret_cast<uint64_t>(out + p.out_off);
                ct.mvx = p.mv_x;
                ct.mvy = p.mv_y;
                ct.pitch = Rk.pitch;
                ct.os = p.out_stride;
                ct.kl = lg;
            }
        }
        const int tasks = ok ? p.blk * p.blk / 16 : 0;
        int incl = tasks;  // inclusive prefix sum over the wavefront's lanes
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = shfl32(incl, lane >= d ? lane - d : lane);
            if (lane >= d) incl += t;
        }
        if (lane == 63) wtot[wave] = incl;
        if (status && ok) status[i] = 0;
        info[tid] = ct;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int v = wtot[k];
            if (k < wave) before += v;
            total += v;
        }
        excl_s[tid] = before + incl - tasks;
        __syncthreads();
        for (int t = tid; t < total; t += 256) {
            int k = 0;  // the last descriptor whose tasks start at or before t
            // (k + sft <= 256 - sft at every step: the read stays in the array)
#pragma unroll
            for (int sft = 128; sft; sft >>= 1)
                if (excl_s[k + sft] <= t) k += sft;
            const ChromaTask c = info[k];
            const int u = t - excl_s[k], kl = c.kl;
            // tile u of the block: (u mod m/4, u div m/4) in units of 4 samples
            chroma_tile(reinterpret_cast<const uint8_t *>(c.y), c.pitch, c.w, c.h, c.clamp != 0,
                        c.mvx, c.mvy,
                        4 * (u & ((1 << (kl - 2)) - 1)), 4 * (u >> (kl - 2)),
                        reinterpret_cast<uint8_t *>(c.o), c.os);
        }
        uint64_t m = __ballot(i < c1 && !ok);
        while (m) {
            const int k = __builtin_ctzll(m);
            m &= m - 1;
            const int64_t j = w0 + 64 * wave + k;
            chroma_block(table, npics, preds[j], j, out, out_bytes, status, R, cached, lane);
        }
        __syncthreads();  // info and excl_s are rewritten by the next round
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
        hipLaunchKernelGGL(mc_luma_kernel,
                           dim3(std::max((n + kMcChunk - 1) / kMcChunk, std::min(n, 8 * device_cus()))),
                           dim3(256), 0, s, table, npics, preds, n, out, out_bytes, status);
    return hipGetLastError();
}

}  // namespace tcme
