// Host runtime and C ABI of libtcme.so (declared in include/tcme.h).
//
// Batch API: a context owns a device picture table; pictures are luma planes
// with their anchor planes; tcme_search() launches the search kernels in the
// context's launch shape (by default one tile-kernel launch for every call).
//
// Drop-in: tc_motion_est() keeps device mirrors of the reference planes it is
// handed.  Before each call it compares, byte for byte, the part of the host
// plane that call can read (the reachable search window plus the six-tap
// apron) with the mirror's host shadow and re-uploads on any difference, so a
// result is always computed from exactly the bytes the caller passed, however
// the caller reuses its buffers.  The call itself is a kernel argument; the
// kernel reads the current block from, and writes the result to, pinned host
// memory, status last, so a call is one launch and a poll of that word.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <thread>
#include <vector>

#include "tcme_internal.h"

#define TCME_API extern "C" __attribute__((visibility("default")))

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(TCME_E_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_),      \
                        __FILE__, __LINE__);                                            \
    } while (0)

struct Pic {
    bool used = false;
    int w = 0, h = 0, pitch = 0, apitch = 0;
    uint8_t *d_y = nullptr;
    uint64_t *d_an = nullptr;  // anchors, apitch * h entries
};

int align_up(int v, int a) { return (v + a - 1) / a * a; }

}  // namespace

struct tcme_ctx {
    int device = 0;
    int max_pics = 0;
    std::vector<Pic> pics;
    tcme::PicDev *d_table = nullptr;
    // staging for tcme_search_host (grown on demand)
    tcme_call *d_calls = nullptr;
    tcme_result *d_res = nullptr;
    tcme_stats *d_stats = nullptr;
    size_t stage_cap = 0;
    // global-motion partial table (grown on demand); successive estimates on
    // one context are ordered by gmv_done, since they share it
    int32_t *d_gmv = nullptr;
    size_t gmv_cap = 0;
    hipEvent_t gmv_done = nullptr;
    bool gmv_pending = false;
    // pinned staging of tcme_pic_upload (grown on demand): a plane is packed
    // there at the device pitch and sent in one copy; up_done orders reuse
    uint8_t *h_up = nullptr;
    size_t up_cap = 0;
    hipEvent_t up_done = nullptr;
    bool up_pending = false;
    std::mutex up_mu;
    tcme::Tiling tiling;
    // plain form: the four per-size kernels run concurrently on these streams,
    // forked from and joined back into the caller's stream with events
    hipStream_t sub[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t fork = nullptr;
    hipEvent_t join[4] = {nullptr, nullptr, nullptr, nullptr};
    // optional per-kernel timing of the last dispatch (tcme_set_timing)
    bool timing = false;
    hipEvent_t t0[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t t1[4] = {nullptr, nullptr, nullptr, nullptr};
    int nlaunched = 0;
    // batched time steps (tcme_tick_search): device layouts per picture shape,
    // the step's call list and picture table, pinned staging of the table
    struct Layout { tcme_call *d = nullptr; int64_t n = 0; };
    std::map<std::tuple<int, int, int, int>, Layout> layouts;
    tcme_call *d_tick_calls = nullptr;
    int64_t tick_cap = 0;
    tcme::TickPicDev *d_tick_pics = nullptr, *h_tick_pics = nullptr;
    int tick_pics_cap = 0;
    hipEvent_t tick_copied = nullptr;
    bool tick_pending = false;
    std::mutex mu;
    std::mutex launch_mu;
};

namespace {

// ---- scratch of the no-allocation entry point ------------------------------
// tcme_global_mv_static keeps its coarse-SAD table in this fixed device array
// instead of a context workspace.  Each GPU the library runs on holds its own
// copy until the process ends.  In host code the array's name is only a
// stand-in: its device address, per device, comes from hipGetSymbolAddress at
// the first tcme_create on that device (scratch_setup) and only that address
// is ever handed to a kernel.
constexpr size_t kScratchBytes = size_t(16) << 20;
}  // namespace
static __device__ __align__(256) int32_t scratch[kScratchBytes / sizeof(int32_t)];
namespace {
constexpr int kMaxDevices = 64;
int32_t *g_scratch_ptr[kMaxDevices];
std::once_flag g_scratch_once[kMaxDevices];
hipError_t g_scratch_err[kMaxDevices];

// Called with `device` current.
int scratch_setup(int device)
{
    if (device < 0 || device >= kMaxDevices) return TCME_OK;  // no scratch there
    std::call_once(g_scratch_once[device], [device] {
        void *p = nullptr;
        g_scratch_err[device] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        g_scratch_ptr[device] = g_scratch_err[device] == hipSuccess ? static_cast<int32_t *>(p) : nullptr;
    });
    if (g_scratch_err[device] != hipSuccess)
        return fail(TCME_E_HIP, "hipGetSymbolAddress(scratch): %s",
                    hipGetErrorString(g_scratch_err[device]));
    return TCME_OK;
}

int set_device(tcme_ctx *c)
{
    HIP_TRY(hipSetDevice(c->device));
    return TCME_OK;
}

tcme::PicDev dev_view(const Pic &p)
{
    tcme::PicDev d{};
    if (!p.used) return d;
    d.y = p.d_y;
    d.an = p.d_an;
    d.w = p.w;
    d.h = p.h;
    d.pitch = p.pitch;
    d.apitch = p.apitch;
    return d;
}

int check_pic(tcme_ctx *c, int pic)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (pic < 0 || pic >= c->max_pics || !c->pics[pic].used)
        return fail(TCME_E_PIC, "picture %d does not exist", pic);
    return TCME_OK;
}

int pic_alloc(tcme_ctx *c, int w, int h, bool half)
{
    if (w <= 0 || h <= 0 || w > 16384 || h > 16384)
        return fail(TCME_E_ARG, "picture size %dx%d out of range", w, h);
    int id = -1;
    for (int i = 0; i < c->max_pics; ++i)
        if (!c->pics[i].used) { id = i; break; }
    if (id < 0) return fail(TCME_E_NOMEM, "picture table full (%d)", c->max_pics);
    Pic p;
    p.w = w;
    p.h = h;
    p.pitch = align_up(w, 64);
    p.apitch = align_up(w, 64);
    size_t ybytes = static_cast<size_t>(p.pitch) * h + tcme::kTailSlack;
    if (hipMalloc(&p.d_y, ybytes) != hipSuccess)
        return fail(TCME_E_NOMEM, "hipMalloc luma %zu bytes", ybytes);
    if (hipMemset(p.d_y, 0, ybytes) != hipSuccess) {
        (void)hipFree(p.d_y);
        return fail(TCME_E_HIP, "clearing luma");
    }
    if (half) {
        size_t abytes = static_cast<size_t>(p.apitch) * h * sizeof(uint64_t) + tcme::kTailSlack;
        if (hipMalloc(&p.d_an, abytes) != hipSuccess) {
            (void)hipFree(p.d_y);
            return fail(TCME_E_NOMEM, "hipMalloc anchor plane %zu bytes", abytes);
        }
        if (hipMemset(p.d_an, 0, abytes) != hipSuccess) {
            (void)hipFree(p.d_y);
            (void)hipFree(p.d_an);
            return fail(TCME_E_HIP, "clearing anchor plane");
        }
    }
    p.used = true;
    c->pics[id] = p;
    tcme::PicDev d = dev_view(p);
    HIP_TRY(hipMemcpy(c->d_table + id, &d, sizeof d, hipMemcpyHostToDevice));
    return id;
}

int pic_free(tcme_ctx *c, int id)
{
    Pic &p = c->pics[id];
    (void)hipFree(p.d_y);
    if (p.d_an) (void)hipFree(p.d_an);
    p = Pic();
    tcme::PicDev d{};
    HIP_TRY(hipMemcpy(c->d_table + id, &d, sizeof d, hipMemcpyHostToDevice));
    return TCME_OK;
}

// Launches the search of n calls in launch shape t on stream s: the kernels
// (the tile kernel for 8x8/16x16 or every size, per-size kernels for the
// rest) run concurrently on the context's streams, forked from and joined
// back into s with events.
int dispatch(tcme_ctx *c, const tcme_call *calls, int n, const tcme::Tiling &t,
             tcme_result *res, tcme_stats *stats, hipStream_t s)
{
    if (n <= 0) return TCME_OK;
    std::lock_guard<std::mutex> g(c->launch_mu);
    static const int sizes[4] = {64, 32, 16, 8};
    static const int slot[4] = {3, 2, 1, 0};
    if (t.tile_waves && t.tile_big) {
        // one launch: the tile kernel takes every call
        if (c->timing) HIP_TRY(hipEventRecord(c->t0[0], s));
        HIP_TRY(tcme::launch_tile(t.tile_waves, true, t.tile_chunk, c->d_table, c->max_pics, calls,
                                  n, res, stats, s));
        if (c->timing) HIP_TRY(hipEventRecord(c->t1[0], s));
        c->nlaunched = 1;
        return TCME_OK;
    }
    HIP_TRY(hipEventRecord(c->fork, s));
    const int nk = t.tile_waves ? 3 : 4;
    for (int k = 0; k < nk; ++k) {
        HIP_TRY(hipStreamWaitEvent(c->sub[k], c->fork, 0));
        if (c->timing) HIP_TRY(hipEventRecord(c->t0[k], c->sub[k]));
        if (t.tile_waves && k == 2)
            HIP_TRY(tcme::launch_tile(t.tile_waves, false, t.tile_chunk, c->d_table, c->max_pics,
                                      calls, n, res, stats, c->sub[k]));
        else
            HIP_TRY(tcme::launch_search_n(sizes[k], t.waves[slot[k]], c->d_table, c->max_pics,
                                          calls, n, res, stats, c->sub[k]));
        if (c->timing) HIP_TRY(hipEventRecord(c->t1[k], c->sub[k]));
        HIP_TRY(hipEventRecord(c->join[k], c->sub[k]));
        HIP_TRY(hipStreamWaitEvent(s, c->join[k], 0));
    }
    c->nlaunched = nk;
    return TCME_OK;
}

int pic_derive(tcme_ctx *c, int id, hipStream_t s)
{
    const Pic &p = c->pics[id];
    if (!p.d_an) return TCME_OK;
    HIP_TRY(tcme::launch_anchors(dev_view(p), s));
    return TCME_OK;
}

}  // namespace

namespace tcme {

int device_cus()
{
    static int cus[64] = {};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (!cus[dev]) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
            v <= 0)
            v = 256;
        cus[dev] = v;
    }
    return cus[dev];
}

}  // namespace tcme

TCME_API const char *tcme_last_error(void) { return g_err.c_str(); }

TCME_API const char *tcme_version(void) { return "tcme 0.1 gfx950"; }

TCME_API int tcme_create(tcme_ctx **out, int device, int max_pics)
{
    if (!out) return fail(TCME_E_ARG, "null out");
    *out = nullptr;
    if (max_pics <= 0) max_pics = 1024;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(TCME_E_HIP, "no HIP device available");
    if (device < 0) HIP_TRY(hipGetDevice(&device));
    if (device >= ndev) return fail(TCME_E_ARG, "device %d >= count %d", device, ndev);
    tcme_ctx *c = new tcme_ctx();
    c->device = device;
    c->max_pics = max_pics;
    c->pics.resize(max_pics);
    if (set_device(c) != TCME_OK) { delete c; return TCME_E_HIP; }
    if (hipMalloc(&c->d_table, sizeof(tcme::PicDev) * max_pics) != hipSuccess) {
        delete c;
        return fail(TCME_E_NOMEM, "picture table");
    }
    bool ok = hipMemset(c->d_table, 0, sizeof(tcme::PicDev) * max_pics) == hipSuccess;
    ok = ok && hipEventCreateWithFlags(&c->fork, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->gmv_done, hipEventDisableTiming) == hipSuccess &&
              hipEventCreateWithFlags(&c->up_done, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; k < 4 && ok; ++k)
        ok = hipStreamCreateWithFlags(&c->sub[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->join[k], hipEventDisableTiming) == hipSuccess &&
             hipEventCreate(&c->t0[k]) == hipSuccess && hipEventCreate(&c->t1[k]) == hipSuccess;
    if (!ok) {
        tcme_destroy(c);
        return fail(TCME_E_HIP, "context streams/events");
    }
    if (int e = scratch_setup(device)) {
        tcme_destroy(c);
        return e;
    }
    *out = c;
    return TCME_OK;
}

TCME_API void tcme_destroy(tcme_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipDeviceSynchronize();
    for (int i = 0; i < c->max_pics; ++i)
        if (c->pics[i].used) pic_free(c, i);
    (void)hipFree(c->d_table);
    (void)hipFree(c->d_calls);
    (void)hipFree(c->d_res);
    (void)hipFree(c->d_stats);
    for (int k = 0; k < 4; ++k) {
        if (c->sub[k]) (void)hipStreamDestroy(c->sub[k]);
        if (c->join[k]) (void)hipEventDestroy(c->join[k]);
        if (c->t0[k]) (void)hipEventDestroy(c->t0[k]);
        if (c->t1[k]) (void)hipEventDestroy(c->t1[k]);
    }
    if (c->fork) (void)hipEventDestroy(c->fork);
    if (c->gmv_done) (void)hipEventDestroy(c->gmv_done);
    if (c->up_done) (void)hipEventDestroy(c->up_done);
    (void)hipFree(c->d_gmv);
    if (c->h_up) (void)hipHostFree(c->h_up);
    for (auto &kv : c->layouts) (void)hipFree(kv.second.d);
    (void)hipFree(c->d_tick_calls);
    (void)hipFree(c->d_tick_pics);
    if (c->h_tick_pics) (void)hipHostFree(c->h_tick_pics);
    if (c->tick_copied) (void)hipEventDestroy(c->tick_copied);
    delete c;
}

TCME_API int tcme_pic_create(tcme_ctx *c, int w, int h)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    return pic_alloc(c, w, h, true);
}

TCME_API int tcme_pic_create_plane(tcme_ctx *c, int w, int h)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    return pic_alloc(c, w, h, false);
}

TCME_API int tcme_pic_destroy(tcme_ctx *c, int pic)
{
    if (int e = check_pic(c, pic)) return e;
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    HIP_TRY(hipDeviceSynchronize());
    return pic_free(c, pic);
}

TCME_API int tcme_pic_upload(tcme_ctx *c, int pic, const uint8_t *host, int stride, void *stream)
{
    if (int e = check_pic(c, pic)) return e;
    const Pic &p = c->pics[pic];
    if (!host || stride < p.w) return fail(TCME_E_ARG, "bad host plane");
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // A 2-D copy from pageable memory goes one row at a time (~50 us a row);
    // packed at the device pitch into pinned staging, the plane goes in one
    // copy.  The caller's buffer is free again on return.
    std::lock_guard<std::mutex> g(c->up_mu);
    if (c->up_pending) {
        HIP_TRY(hipEventSynchronize(c->up_done));
        c->up_pending = false;
    }
    const size_t bytes = static_cast<size_t>(p.pitch) * p.h;
    if (bytes > c->up_cap) {
        if (c->h_up) HIP_TRY(hipHostFree(c->h_up));
        c->h_up = nullptr;
        c->up_cap = 0;
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&c->h_up), bytes, 0));
        c->up_cap = bytes;
    }
    for (int y = 0; y < p.h; ++y) {
        uint8_t *row = c->h_up + static_cast<size_t>(y) * p.pitch;
        memcpy(row, host + static_cast<size_t>(y) * stride, p.w);
        memset(row + p.w, 0, p.pitch - p.w);  // the row padding stays zero
    }
    HIP_TRY(hipMemcpyAsync(p.d_y, c->h_up, bytes, hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(c->up_done, s));
    c->up_pending = true;
    return pic_derive(c, pic, s);
}

TCME_API int tcme_pic_upload_device(tcme_ctx *c, int pic, const uint8_t *dev, int stride,
                                    void *stream)
{
    if (int e = check_pic(c, pic)) return e;
    const Pic &p = c->pics[pic];
    if (!dev || stride < p.w) return fail(TCME_E_ARG, "bad device plane");
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    HIP_TRY(hipMemcpy2DAsync(p.d_y, p.pitch, dev, stride, p.w, p.h, hipMemcpyDeviceToDevice, s));
    return pic_derive(c, pic, s);
}

TCME_API int tcme_pic_luma(tcme_ctx *c, int pic, uint8_t **dev, int *pitch)
{
    if (int e = check_pic(c, pic)) return e;
    if (!dev || !pitch) return fail(TCME_E_ARG, "null output");
    *dev = c->pics[pic].d_y;
    *pitch = c->pics[pic].pitch;
    return TCME_OK;
}

TCME_API int tcme_pic_finish(tcme_ctx *c, int pic, void *stream)
{
    if (int e = check_pic(c, pic)) return e;
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    return pic_derive(c, pic, static_cast<hipStream_t>(stream));
}

TCME_API int tcme_search(tcme_ctx *c, const tcme_call *calls, int n, tcme_result *results,
                         tcme_stats *stats, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!calls || !results))) return fail(TCME_E_ARG, "bad call list");
    if (n == 0) return TCME_OK;
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    return dispatch(c, calls, n, c->tiling, results, stats, static_cast<hipStream_t>(stream));
}

namespace {

int check_tiling(const tcme_tiling *t)
{
    if (!t) return fail(TCME_E_ARG, "null tiling");
    const int w[4] = {t->waves8, t->waves16, t->waves32, t->waves64};
    for (int k = 0; k < 4; ++k)
        if (!tcme::valid_waves(8 << k, w[k]))
            return fail(TCME_E_ARG, "%d waves per %dx%d call not supported", w[k], 8 << k, 8 << k);
    if (t->tile_waves != 0 && t->tile_waves != 4 && t->tile_waves != 8)
        return fail(TCME_E_ARG, "%d waves per tile workgroup not supported", t->tile_waves);
    if (t->tile_waves && (t->tile_chunk < 1 || t->tile_chunk > tcme::kTileChunkMax))
        return fail(TCME_E_ARG, "tile chunk %d outside 1..%d", t->tile_chunk, tcme::kTileChunkMax);
    if (t->tile_big != 0 && !(t->tile_big == 1 && (t->tile_waves == 4 || t->tile_waves == 8)))
        return fail(TCME_E_ARG, "tile_big needs tile_waves 4 or 8");
    return TCME_OK;
}

tcme::Tiling to_internal(const tcme_tiling &t)
{
    tcme::Tiling r;
    r.waves[0] = t.waves8;
    r.waves[1] = t.waves16;
    r.waves[2] = t.waves32;
    r.waves[3] = t.waves64;
    r.tile_waves = t.tile_waves;
    r.tile_chunk = t.tile_waves ? t.tile_chunk : r.tile_chunk;
    r.tile_big = t.tile_big;
    return r;
}

}  // namespace

TCME_API int tcme_set_tiling(tcme_ctx *c, const tcme_tiling *t)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (int e = check_tiling(t)) return e;
    c->tiling = to_internal(*t);
    return TCME_OK;
}

TCME_API int tcme_get_tiling(tcme_ctx *c, tcme_tiling *t)
{
    if (!c || !t) return fail(TCME_E_ARG, "null argument");
    t->waves8 = c->tiling.waves[0];
    t->waves16 = c->tiling.waves[1];
    t->waves32 = c->tiling.waves[2];
    t->waves64 = c->tiling.waves[3];
    t->tile_waves = c->tiling.tile_waves;
    t->tile_chunk = c->tiling.tile_chunk;
    t->tile_big = c->tiling.tile_big;
    return TCME_OK;
}

// Times every distinct launch shape (tile workgroups of 4 or 8 wavefronts for
// every call; or tile workgroups of 0, 4 or 8 wavefronts for the 8x8/16x16
// calls with each count of wavefronts per 32x32 and per 64x64 call; the tile
// chunk stays the context's) on
// the caller's call list (device arrays; results are overwritten with the
// same answers) and keeps the fastest in the context.  Results do not depend
// on the shape, only time does.
TCME_API int tcme_autotune(tcme_ctx *c, const tcme_call *calls, int n, tcme_result *results,
                           int reps, void *stream, tcme_tiling *best_out, float *best_ms)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n <= 0 || !calls || !results) return fail(TCME_E_ARG, "bad call list");
    if (reps < 1) reps = 3;
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // the distinct shapes: tile_big workgroups of 4 or 8 wavefronts (the
    // per-call counts unused), then tile workgroups of 0, 4 or 8 wavefronts
    // for the small calls with every per-call count for 32x32 and 64x64
    std::vector<tcme::Tiling> shapes;
    for (int tw : {8, 4}) {
        tcme::Tiling t;
        t.tile_waves = tw;
        t.tile_big = 1;
        shapes.push_back(t);
    }
    for (int tw : {0, 4, 8})
        for (int a : {1, 2, 4})
            for (int b : {2, 4, 8}) {
                tcme::Tiling t;
                t.waves[2] = a;
                t.waves[3] = b;
                t.tile_waves = tw;
                t.tile_big = 0;
                shapes.push_back(t);
            }
    for (tcme::Tiling &t : shapes) t.tile_chunk = c->tiling.tile_chunk;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int err = TCME_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        err = fail(TCME_E_HIP, "timing events");
    tcme::Tiling best = c->tiling;
    float best_t = 1e30f;
    for (size_t k = 0; k < shapes.size() && !err; ++k) {
        const tcme::Tiling &t = shapes[k];
        err = dispatch(c, calls, n, t, results, nullptr, s);  // warm
        if (!err && hipEventRecord(e0, s) != hipSuccess) err = fail(TCME_E_HIP, "event record");
        for (int r = 0; r < reps && !err; ++r) err = dispatch(c, calls, n, t, results, nullptr, s);
        float ms = 0;
        if (!err && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                     hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
            err = fail(TCME_E_HIP, "event timing");
        if (!err && ms / reps < best_t) {
            best_t = ms / reps;
            best = t;
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (err) return err;
    c->tiling = best;
    if (best_out) tcme_get_tiling(c, best_out);
    if (best_ms) *best_ms = best_t;
    return TCME_OK;
}

struct tcme_graph {
    tcme_ctx *ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

// Captures dispatch() on a private stream: the fork/join events pull the
// context's side streams into the capture, so the graph holds the same
// concurrent kernels.  dispatch() makes no allocation, copy or memset.
TCME_API int tcme_graph_create(tcme_ctx *c, const tcme_call *calls, int n, tcme_result *results,
                               tcme_stats *stats, tcme_graph **out)
{
    if (!c || !out) return fail(TCME_E_ARG, "null argument");
    *out = nullptr;
    if (n <= 0 || !calls || !results) return fail(TCME_E_ARG, "bad call list");
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t cs = nullptr;
    HIP_TRY(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    const bool timing = c->timing;
    c->timing = false;
    hipGraph_t graph = nullptr;
    int err = TCME_OK;
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        err = fail(TCME_E_HIP, "begin capture");
    } else {
        err = dispatch(c, calls, n, c->tiling, results, stats, cs);
        const hipError_t e = hipStreamEndCapture(cs, &graph);
        if (!err && e != hipSuccess) err = fail(TCME_E_HIP, "end capture: %s", hipGetErrorString(e));
    }
    c->timing = timing;
    (void)hipStreamDestroy(cs);
    hipGraphExec_t exec = nullptr;
    if (!err && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
        err = fail(TCME_E_HIP, "graph instantiate");
    if (err) {
        if (graph) (void)hipGraphDestroy(graph);
        return err;
    }
    *out = new tcme_graph{c, graph, exec};
    return TCME_OK;
}

TCME_API int tcme_graph_launch(tcme_graph *g, void *stream)
{
    if (!g) return fail(TCME_E_ARG, "null graph");
    if (set_device(g->ctx) != TCME_OK) return TCME_E_HIP;
    HIP_TRY(hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)));
    return TCME_OK;
}

TCME_API void tcme_graph_destroy(tcme_graph *g)
{
    if (!g) return;
    (void)hipSetDevice(g->ctx->device);
    (void)hipDeviceSynchronize();
    (void)hipGraphExecDestroy(g->exec);
    (void)hipGraphDestroy(g->graph);
    delete g;
}

namespace {

int predict(tcme_ctx *c, bool chroma, const tcme_pred *preds, int n, uint8_t *out,
            int64_t out_bytes, int32_t *status, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!preds || !out || out_bytes <= 0)))
        return fail(TCME_E_ARG, "bad prediction list");
    if (n == 0) return TCME_OK;
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    HIP_TRY(tcme::launch_predict(chroma, c->d_table, c->max_pics, preds, n, out, out_bytes, status,
                                 static_cast<hipStream_t>(stream)));
    return TCME_OK;
}

}  // namespace

TCME_API int tcme_predict_luma(tcme_ctx *c, const tcme_pred *preds, int n, uint8_t *out,
                               int64_t out_bytes, int32_t *status, void *stream)
{
    return predict(c, false, preds, n, out, out_bytes, status, stream);
}

TCME_API int tcme_predict_chroma(tcme_ctx *c, const tcme_pred *preds, int n, uint8_t *out,
                                 int64_t out_bytes, int32_t *status, void *stream)
{
    return predict(c, true, preds, n, out, out_bytes, status, stream);
}

TCME_API int tcme_global_mv(tcme_ctx *c, const int32_t *pairs, int n, int32_t *out, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!pairs || !out))) return fail(TCME_E_ARG, "bad pair list");
    if (n > 65535) return fail(TCME_E_ARG, "%d pairs in one call (at most 65535)", n);
    if (n == 0) return TCME_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    int max_w = 1, max_h = 1;
    for (const Pic &p : c->pics)
        if (p.used) { max_w = std::max(max_w, p.w); max_h = std::max(max_h, p.h); }
    const size_t need = tcme::gmv_part_bytes(n, max_w, max_h);
    if (need > c->gmv_cap) {
        if (c->gmv_pending) HIP_TRY(hipEventSynchronize(c->gmv_done));
        HIP_TRY(hipFree(c->d_gmv));
        c->d_gmv = nullptr;
        c->gmv_cap = 0;
        if (hipMalloc(&c->d_gmv, need) != hipSuccess)
            return fail(TCME_E_NOMEM, "global-motion table %zu bytes", need);
        c->gmv_cap = need;
    } else if (c->gmv_pending) {
        HIP_TRY(hipStreamWaitEvent(s, c->gmv_done, 0));
    }
    HIP_TRY(tcme::launch_global_mv(c->d_table, c->max_pics, pairs, n, max_w, max_h, c->d_gmv, out,
                                   s));
    HIP_TRY(hipEventRecord(c->gmv_done, s));
    c->gmv_pending = true;
    return TCME_OK;
}

// tcme_global_mv without allocation: the coarse-SAD table lives in the fixed
// per-device scratch array.  Every entry the walk reads is written by the
// table kernel first, so nothing needs clearing between calls.
TCME_API int tcme_global_mv_static(tcme_ctx *c, const int32_t *pairs, int n, int32_t *out,
                                   void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!pairs || !out))) return fail(TCME_E_ARG, "bad pair list");
    if (n > 65535) return fail(TCME_E_ARG, "%d pairs in one call (at most 65535)", n);
    if (n == 0) return TCME_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    if (c->device < 0 || c->device >= kMaxDevices || !g_scratch_ptr[c->device])
        return fail(TCME_E_HIP, "no scratch address for device %d", c->device);
    hipStream_t s = static_cast<hipStream_t>(stream);
    int max_w = 1, max_h = 1;
    for (const Pic &p : c->pics)
        if (p.used) { max_w = std::max(max_w, p.w); max_h = std::max(max_h, p.h); }
    const size_t need = tcme::gmv_part_bytes(n, max_w, max_h);
    if (need > kScratchBytes)
        return fail(TCME_E_ARG, "%d pairs of pictures up to %dx%d need %zu bytes of scratch "
                    "(at most %zu): use fewer pairs per call, or tcme_global_mv",
                    n, max_w, max_h, need, kScratchBytes);
    if (c->gmv_pending) HIP_TRY(hipStreamWaitEvent(s, c->gmv_done, 0));
    HIP_TRY(tcme::launch_global_mv(c->d_table, c->max_pics, pairs, n, max_w, max_h,
                                   g_scratch_ptr[c->device], out, s));
    HIP_TRY(hipEventRecord(c->gmv_done, s));
    c->gmv_pending = true;
    return TCME_OK;
}

TCME_API int tcme_sad(tcme_ctx *c, const tcme_sad_desc *descs, int n, const uint8_t *buf,
                      int64_t buf_bytes, int32_t *out, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!descs || !out))) return fail(TCME_E_ARG, "bad SAD list");
    if (n == 0) return TCME_OK;
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    HIP_TRY(tcme::launch_sad(c->d_table, c->max_pics, descs, n, buf, buf ? buf_bytes : 0, out,
                             static_cast<hipStream_t>(stream)));
    return TCME_OK;
}

TCME_API int tcme_search_chain(tcme_ctx *c, const tcme_chain_frame *frames, int n,
                              tcme_chain_call *calls, int32_t *counts, int64_t ctu_cap,
                              void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 1 || n > tcme::kChainFrames || !frames || !calls || !counts)
        return fail(TCME_E_ARG, "bad chain frame list (1..%d frames)", tcme::kChainFrames);
    tcme::ChainJobHost jobs[tcme::kChainFrames];
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const tcme_chain_frame &f = frames[i];
        if (int e = check_pic(c, f.cur_pic)) return e;
        const Pic &cp = c->pics[f.cur_pic];
        if (f.nrefs < 1 || f.nrefs > 4 || f.sr < 0 || f.sr > 16383 || f.lambda < 0 ||
            f.lambda > (1 << 24))
            return fail(TCME_E_ARG, "chain frame %d: nrefs %d, sr %d, lambda %d", i, f.nrefs,
                        f.sr, f.lambda);
        for (int r = 0; r < f.nrefs; ++r) {
            if (int e = check_pic(c, f.ref_pic[r])) return e;
            const Pic &rp = c->pics[f.ref_pic[r]];
            if (!rp.d_an || rp.w != cp.w || rp.h != cp.h)
                return fail(TCME_E_PIC, "chain frame %d: reference %d is not a %dx%d reference "
                            "picture", i, r, cp.w, cp.h);
        }
        tcme::ChainJobHost &j = jobs[i];
        j.cur_pic = f.cur_pic;
        j.nrefs = f.nrefs;
        for (int r = 0; r < 4; ++r) j.ref_pic[r] = r < f.nrefs ? f.ref_pic[r] : f.ref_pic[0];
        j.sr = f.sr;
        j.lambda = f.lambda;
        j.first = static_cast<int>(total);
        j.cols = (cp.w + 63) / 64;
        total += static_cast<int64_t>(j.cols) * ((cp.h + 63) / 64);
    }
    if (total > ctu_cap) return fail(TCME_E_ARG, "%lld CTUs, room for %lld", (long long)total,
                                     (long long)ctu_cap);
    if (total > (1 << 30)) return fail(TCME_E_ARG, "too many CTUs");
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    HIP_TRY(tcme::launch_chain(c->d_table, jobs, n, static_cast<int>(total), calls, counts,
                               static_cast<hipStream_t>(stream)));
    return TCME_OK;
}

namespace {

// The calls of one picture shape as a time step lays them out: every
// in-frame quadtree node (64, 32, 16, 8) against each reference, centred at
// the block position; complete CTUs (all 85 nodes in the frame) first, then
// the partial ones at the edges; reference-major; CTUs in raster order; in a
// CTU the largest blocks first, each size in raster order.  ref_pic holds the
// reference index, cur_pic -1.
std::vector<tcme_call> make_layout(int w, int h, int nrefs, int sr)
{
    const int cols = (w + 63) / 64, rows = (h + 63) / 64;
    std::vector<std::vector<std::pair<int, int>>> nodes(static_cast<size_t>(cols) * rows);
    std::vector<std::vector<int>> sizes(nodes.size());
    for (int n = 64; n >= 8; n >>= 1)
        for (int y = 0; y + n <= h; y += n)
            for (int x = 0; x + n <= w; x += n) {
                const size_t k = static_cast<size_t>(y / 64) * cols + x / 64;
                nodes[k].emplace_back(x, y);
                sizes[k].push_back(n);
            }
    std::vector<tcme_call> out;
    for (int partial = 0; partial < 2; ++partial)
        for (int r = 0; r < nrefs; ++r)
            for (size_t k = 0; k < nodes.size(); ++k) {
                if ((nodes[k].size() != 85) != (partial == 1)) continue;
                for (size_t i = 0; i < nodes[k].size(); ++i) {
                    const int x = nodes[k][i].first, y = nodes[k][i].second, n = sizes[k][i];
                    out.push_back(tcme_call{r, -1, x, y, x, y, n, sr});
                }
            }
    return out;
}

}  // namespace

TCME_API int64_t tcme_tick_layout(tcme_ctx *c, int w, int h, int nrefs, int sr, tcme_call *calls,
                                  int64_t cap)
{
    (void)c;  // the layout depends on the shape only (no device work: c may be NULL)
    if (w <= 0 || h <= 0 || w > tcme::kMaxPicSide || h > tcme::kMaxPicSide || nrefs < 1 ||
        nrefs > 4 || sr < 0 || sr > 16383)
        return fail(TCME_E_ARG, "layout %dx%d, %d references, range %d", w, h, nrefs, sr);
    const std::vector<tcme_call> v = make_layout(w, h, nrefs, sr);
    if (calls) {
        if (cap < static_cast<int64_t>(v.size())) return fail(TCME_E_ARG, "room for %lld of %zu calls",
                                                             (long long)cap, v.size());
        memcpy(calls, v.data(), v.size() * sizeof(tcme_call));
    }
    return static_cast<int64_t>(v.size());
}

TCME_API int tcme_tick_search(tcme_ctx *c, const tcme_tick_pic *pics, int n, tcme_result *results,
                              int64_t results_cap, int64_t *offsets, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!pics || !results || !offsets))) return fail(TCME_E_ARG, "bad tick");
    if (offsets) offsets[0] = 0;
    if (n == 0) return TCME_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // validate, find (or build) each picture's layout
    std::vector<tcme::TickPicDev> tab(n);
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const tcme_tick_pic &p = pics[i];
        if (p.cur_pic < 0 || p.cur_pic >= c->max_pics || !c->pics[p.cur_pic].used)
            return fail(TCME_E_PIC, "tick picture %d: no picture %d", i, p.cur_pic);
        const Pic &cp = c->pics[p.cur_pic];
        if (p.nrefs < 1 || p.nrefs > 4 || p.sr < 0 || p.sr > 16383)
            return fail(TCME_E_ARG, "tick picture %d: %d references, range %d", i, p.nrefs, p.sr);
        tcme::TickPicDev &t = tab[i];
        for (int r = 0; r < 4; ++r) {
            const int id = r < p.nrefs ? p.ref_pic[r] : p.ref_pic[0];
            if (id < 0 || id >= c->max_pics || !c->pics[id].used || !c->pics[id].d_an ||
                c->pics[id].w != cp.w || c->pics[id].h != cp.h)
                return fail(TCME_E_PIC, "tick picture %d: reference %d is not a %dx%d reference "
                            "picture", i, r, cp.w, cp.h);
            t.ref_pic[r] = id;
        }
        const auto key = std::make_tuple(cp.w, cp.h, p.nrefs, p.sr);
        auto it = c->layouts.find(key);
        if (it == c->layouts.end()) {
            const std::vector<tcme_call> v = make_layout(cp.w, cp.h, p.nrefs, p.sr);
            tcme_ctx::Layout L;
            L.n = static_cast<int64_t>(v.size());
            if (L.n > 0) {
                HIP_TRY(hipStreamSynchronize(s));
                if (hipMalloc(&L.d, v.size() * sizeof(tcme_call)) != hipSuccess)
                    return fail(TCME_E_NOMEM, "layout of %lld calls", (long long)L.n);
                HIP_TRY(hipMemcpy(L.d, v.data(), v.size() * sizeof(tcme_call), hipMemcpyHostToDevice));
            }
            it = c->layouts.emplace(key, L).first;
        }
        t.tmpl = it->second.d;
        t.ncalls = static_cast<int32_t>(it->second.n);
        t.first = total;
        t.cur_pic = p.cur_pic;
        total += it->second.n;
        offsets[i + 1] = total;
    }
    if (total > results_cap)
        return fail(TCME_E_ARG, "%lld calls, room for %lld results", (long long)total,
                    (long long)results_cap);
    if (total > INT32_MAX) return fail(TCME_E_ARG, "%lld calls in one step", (long long)total);
    // the step's call list and picture table (grown on demand; the previous
    // table copy must be done before its pinned staging is rewritten)
    if (c->tick_pending) {
        HIP_TRY(hipEventSynchronize(c->tick_copied));
        c->tick_pending = false;
    }
    if (total > c->tick_cap) {
        HIP_TRY(hipStreamSynchronize(s));
        (void)hipFree(c->d_tick_calls);
        c->d_tick_calls = nullptr;
        c->tick_cap = 0;
        if (hipMalloc(&c->d_tick_calls, total * sizeof(tcme_call)) != hipSuccess)
            return fail(TCME_E_NOMEM, "step call list of %lld calls", (long long)total);
        c->tick_cap = total;
    }
    if (n > c->tick_pics_cap) {
        HIP_TRY(hipStreamSynchronize(s));
        (void)hipFree(c->d_tick_pics);
        if (c->h_tick_pics) (void)hipHostFree(c->h_tick_pics);
        c->d_tick_pics = nullptr;
        c->h_tick_pics = nullptr;
        c->tick_pics_cap = 0;
        const int cap = std::max(n, 256);
        if (hipMalloc(&c->d_tick_pics, cap * sizeof(tcme::TickPicDev)) != hipSuccess ||
            hipHostMalloc(reinterpret_cast<void **>(&c->h_tick_pics), cap * sizeof(tcme::TickPicDev), 0) !=
                hipSuccess)
            return fail(TCME_E_NOMEM, "step picture table");
        c->tick_pics_cap = cap;
    }
    if (!c->tick_copied) HIP_TRY(hipEventCreateWithFlags(&c->tick_copied, hipEventDisableTiming));
    memcpy(c->h_tick_pics, tab.data(), n * sizeof(tcme::TickPicDev));
    HIP_TRY(hipMemcpyAsync(c->d_tick_pics, c->h_tick_pics, n * sizeof(tcme::TickPicDev),
                           hipMemcpyHostToDevice, s));
    HIP_TRY(hipEventRecord(c->tick_copied, s));
    c->tick_pending = true;
    HIP_TRY(tcme::launch_expand(c->d_tick_pics, n, c->d_tick_calls, total, s));
    return dispatch(c, c->d_tick_calls, static_cast<int>(total), c->tiling, results, nullptr, s);
}

TCME_API int tcme_set_timing(tcme_ctx *c, int on)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    c->timing = on != 0;
    c->nlaunched = 0;
    return TCME_OK;
}

TCME_API int tcme_last_timing(tcme_ctx *c, float *ms, int max)
{
    if (!c || (max > 0 && !ms)) return fail(TCME_E_ARG, "null argument");
    if (!c->timing) return fail(TCME_E_ARG, "timing is off");
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    const int n = c->nlaunched;
    for (int k = 0; k < n && k < max; ++k) {
        HIP_TRY(hipEventSynchronize(c->t1[k]));
        HIP_TRY(hipEventElapsedTime(&ms[k], c->t0[k], c->t1[k]));
    }
    return n;
}

TCME_API int tcme_search_host(tcme_ctx *c, const tcme_call *calls, int n, tcme_result *results,
                              tcme_stats *stats, void *stream)
{
    if (!c) return fail(TCME_E_ARG, "null context");
    if (n < 0 || (n > 0 && (!calls || !results))) return fail(TCME_E_ARG, "bad call list");
    if (n == 0) return TCME_OK;
    std::lock_guard<std::mutex> g(c->mu);
    if (set_device(c) != TCME_OK) return TCME_E_HIP;
    hipStream_t s = static_cast<hipStream_t>(stream);
    if (static_cast<size_t>(n) > c->stage_cap) {
        HIP_TRY(hipStreamSynchronize(s));
        (void)hipFree(c->d_calls);
        (void)hipFree(c->d_res);
        (void)hipFree(c->d_stats);
        c->d_calls = nullptr; c->d_res = nullptr; c->d_stats = nullptr; c->stage_cap = 0;
        size_t cap = std::max<size_t>(n, 4096);
        if (hipMalloc(&c->d_calls, cap * sizeof(tcme_call)) != hipSuccess ||
            hipMalloc(&c->d_res, cap * sizeof(tcme_result)) != hipSuccess ||
            hipMalloc(&c->d_stats, cap * sizeof(tcme_stats)) != hipSuccess)
            return fail(TCME_E_NOMEM, "staging for %d calls", n);
        c->stage_cap = cap;
    }
    HIP_TRY(hipMemcpyAsync(c->d_calls, calls, n * sizeof(tcme_call), hipMemcpyHostToDevice, s));
    if (int e = dispatch(c, c->d_calls, n, c->tiling, c->d_res, stats ? c->d_stats : nullptr, s))
        return e;
    HIP_TRY(hipMemcpyAsync(results, c->d_res, n * sizeof(tcme_result), hipMemcpyDeviceToHost, s));
    if (stats)
        HIP_TRY(hipMemcpyAsync(stats, c->d_stats, n * sizeof(tcme_stats), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return TCME_OK;
}

// ---------------------------------------------------------------------------
// Drop-in tc_motion_est
// ---------------------------------------------------------------------------
namespace {

struct Mirror {
    const uint8_t *host = nullptr;
    int w = 0, h = 0, stride = 0;
    int pic = -1;
    std::vector<uint8_t> shadow;  // w*h copy of what the device holds
    uint64_t last_use = 0;
};

struct DropIn {
    std::mutex mu;
    tcme_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;
    std::vector<Mirror> mirrors;
    uint64_t tick = 0;
    // Pinned, coherent host memory that the drop-in kernel reads and writes
    // directly: the current block (rows packed at pitch blk) and the result.
    uint8_t *h_blk = nullptr;
    tcme_result *h_res = nullptr;
    const uint8_t *d_blk = nullptr;   // device addresses of the two
    tcme_result *d_res = nullptr;
    size_t blk_cap = 0;               // bytes of h_blk
    long calls = 0, uploads = 0;
};

DropIn g_drop;
constexpr int kMirrors = 8;
std::once_flag g_drop_once;
std::string g_drop_init_err;  // set by a failed start (dropin_init_body)

[[noreturn]] void die(const char *what)
{
    fprintf(stderr, "tcme: fatal: %s: %s\n", what, tcme_last_error());
    abort();
}

void dropin_init_body();

void dropin_ensure()
{
    std::call_once(g_drop_once, dropin_init_body);
    if (!g_drop_init_err.empty()) {
        fail(TCME_E_HIP, "%s", g_drop_init_err.c_str());
        die("drop-in");
    }
}

// The drop-in's device context, pinned staging and stream.  Run once (it may
// run on tcme_prepare's thread); failures are kept in init_err and reported
// by the first call that needs the GPU (dropin_ensure).
void dropin_init_body()
{
    DropIn &d = g_drop;
    const char *dev = getenv("TCME_DEVICE");
    if (tcme_create(&d.ctx, dev ? atoi(dev) : 0, 64) != TCME_OK) {
        g_drop_init_err = std::string("no usable GPU: ") + tcme_last_error();
        d.ctx = nullptr;
        return;
    }
    if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess) {
        g_drop_init_err = "drop-in stream";
        return;
    }
    // The kernel reads each 4-byte unit of the block as the aligned dword pair
    // that covers it, so the last unit of a 64x64 block also reads bytes
    // 4096-4099: the buffer carries the same tail slack as a device picture.
    void *db = nullptr, *dr = nullptr;
    if (hipHostMalloc(reinterpret_cast<void **>(&d.h_blk),
                      tcme::kMaxBlk * tcme::kMaxBlk + tcme::kTailSlack,
                      hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&d.h_res), sizeof(tcme_result),
                      hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(&db, d.h_blk, 0) != hipSuccess ||
        hipHostGetDevicePointer(&dr, d.h_res, 0) != hipSuccess) {
        g_drop_init_err = "drop-in staging";
        return;
    }
    d.d_blk = static_cast<const uint8_t *>(db);
    d.d_res = static_cast<tcme_result *>(dr);
    d.blk_cap = tcme::kMaxBlk * tcme::kMaxBlk + tcme::kTailSlack;
}

// Device picture holding the bytes of host plane `ref` that the call reading
// rows [y0,y1] x cols [x0,x1] depends on.
int mirror_for(const uint8_t *ref, int stride, int w, int h, int x0, int y0, int x1, int y1)
{
    DropIn &d = g_drop;
    Mirror *m = nullptr;
    for (auto &e : d.mirrors)
        if (e.host == ref && e.w == w && e.h == h && e.stride == stride) { m = &e; break; }
    bool stale = true;
    if (m) {
        stale = false;
        const size_t span = static_cast<size_t>(x1 - x0 + 1);
        for (int y = y0; y <= y1 && !stale; ++y)
            stale = memcmp(m->shadow.data() + static_cast<size_t>(y) * w + x0,
                           ref + static_cast<size_t>(y) * stride + x0, span) != 0;
    } else {
        if (static_cast<int>(d.mirrors.size()) < kMirrors) {
            d.mirrors.emplace_back();
            m = &d.mirrors.back();
        } else {
            m = &*std::min_element(d.mirrors.begin(), d.mirrors.end(),
                                   [](const Mirror &a, const Mirror &b) { return a.last_use < b.last_use; });
        }
        std::lock_guard<std::mutex> gc(d.ctx->mu);  // the picture table is shared
        if (m->pic >= 0 && (m->w != w || m->h != h)) {
            if (hipStreamSynchronize(d.stream) != hipSuccess) die("stream synchronize");
            pic_free(d.ctx, m->pic);
            m->pic = -1;
        }
        if (m->pic < 0) {
            m->pic = pic_alloc(d.ctx, w, h, true);
            if (m->pic < 0) die("reference mirror");
        }
        m->host = ref;
        m->w = w;
        m->h = h;
        m->stride = stride;
    }
    if (stale) {
        m->shadow.resize(static_cast<size_t>(w) * h);
        for (int y = 0; y < h; ++y)
            memcpy(m->shadow.data() + static_cast<size_t>(y) * w, ref + static_cast<size_t>(y) * stride, w);
        if (tcme_pic_upload(d.ctx, m->pic, m->shadow.data(), w, d.stream) != TCME_OK) die("upload");
        d.uploads++;
    }
    m->last_use = ++d.tick;
    return m->pic;
}

}  // namespace

TCME_API tcme_mv tc_motion_est(const uint8_t *ref, int ref_stride, int ref_w, int ref_h,
                               const uint8_t *orig, int orig_stride, int center_x, int center_y,
                               int blk_size, int search_range, int32_t *best_sad_out)
{
    DropIn &d = g_drop;
    dropin_ensure();
    std::lock_guard<std::mutex> g(d.mu);
    if (blk_size < 1 || blk_size > tcme::kMaxPicSide) {
        fail(TCME_E_BLK, "block size %d", blk_size);
        die("tc_motion_est");
    }
    if (ref_w <= 0 || ref_h <= 0 || ref_stride < ref_w) {
        fail(TCME_E_ARG, "reference plane %dx%d stride %d", ref_w, ref_h, ref_stride);
        die("tc_motion_est");
    }
    const int n = blk_size;
    // Reachable window: two large-hex passes (16*scale), three small-hex
    // passes (12), four diamond passes (4), the quarter-pel floor (1) and the
    // six-tap apron (2 left/up, 3 right/down) - margin 16*scale + 24.
    // (the step is capped at the larger side, as the kernels do: beyond that
    // every large-hex point is off the picture and the window is the whole
    // picture anyway)
    const int scale = tcme::search_scale(search_range, ref_w, ref_h);
    const int reach = 16 * scale + 24;
    int cx = center_x < 0 ? 0 : (center_x > ref_w - n ? ref_w - n : center_x);
    int cy = center_y < 0 ? 0 : (center_y > ref_h - n ? ref_h - n : center_y);
    const int x0 = std::max(0, cx - reach), y0 = std::max(0, cy - reach);
    const int x1 = std::min(ref_w - 1, cx + n - 1 + reach), y1 = std::min(ref_h - 1, cy + n - 1 + reach);
    if (hipSetDevice(d.ctx->device) != hipSuccess) die("hipSetDevice");
    const int rpic = (x0 <= x1 && y0 <= y1) ? mirror_for(ref, ref_stride, ref_w, ref_h, x0, y0, x1, y1)
                                            : mirror_for(ref, ref_stride, ref_w, ref_h, 0, 0, -1, -1);

    // The call goes as a kernel argument (cur_pic and px/py are unused: the
    // kernel reads the block from h_blk); the kernel overwrites the sentinel
    // status, so a launch that did not run cannot pass for a result.
    const tcme_call call = {rpic, -1, 0, 0, center_x, center_y, n, search_range};
    const size_t need = static_cast<size_t>(n) * n + tcme::kTailSlack;
    if (need > d.blk_cap) {
        // a block larger than 64x64 (only the generic kernel takes those)
        if (hipStreamSynchronize(d.stream) != hipSuccess) die("stream synchronize");
        (void)hipHostFree(d.h_blk);
        void *db = nullptr;
        if (hipHostMalloc(reinterpret_cast<void **>(&d.h_blk), need, hipHostMallocCoherent) != hipSuccess ||
            hipHostGetDevicePointer(&db, d.h_blk, 0) != hipSuccess)
            die("drop-in block staging");
        d.d_blk = static_cast<const uint8_t *>(db);
        d.blk_cap = need;
    }
    for (int r = 0; r < n; ++r) memcpy(d.h_blk + r * n, orig + static_cast<size_t>(r) * orig_stride, n);
    volatile tcme_result *hr = d.h_res;
    hr->status = TCME_E_HIP;
    if (tcme::launch_dropin(d.ctx->d_table, call, d.d_blk, d.d_res, d.stream) != hipSuccess) die("launch");
    // The kernel writes the status last, after a system-scope fence: poll it
    // (a blocking synchronisation adds a wake-up to every call).  Past 2 ms
    // without it, synchronise, which also reports a failed launch.  The kernel
    // reads the block only before it writes the result, so the next call may
    // refill it while this kernel is still retiring.
    const auto t0 = std::chrono::steady_clock::now();
    bool seen = false;
    for (unsigned spin = 1;; ++spin) {
        if (hr->status != TCME_E_HIP) { seen = true; break; }
        if ((spin & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2))
            break;
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    if (!seen && hipStreamSynchronize(d.stream) != hipSuccess) die("result");
    if (hr->status != 0) {
        fail(hr->status, "kernel rejected the call (status %d)", hr->status);
        die("tc_motion_est");
    }
    d.calls++;
    if (best_sad_out) *best_sad_out = hr->sad;
    tcme_mv mv = {hr->mv_x, hr->mv_y};
    return mv;
}

// ---------------------------------------------------------------------------
// Frame-level drop-in (tcme_frame_*)
//
// An encoder declares each frame's planes once (tcme_frame_begin) together
// with the searches its call sites will make at centres known at frame start.
// Those are searched at once, in one launch, and every later call with
// exactly the same arguments is answered from their results.  A call that
// finds no result is searched at once on the resident planes, in one launch
// together with the other nodes of its CTU at the same displacement (the
// reference's predictor is CTU-local, so they are the likely next calls).
// Results are pure functions of the arguments, so the answers are the
// reference's whichever way they were found.
// ---------------------------------------------------------------------------
namespace {

constexpr int kFrameRefs = 8;       // planes a frame may declare
constexpr int kRing = 1 << 15;      // miss-batch result slots
constexpr int kBatchMax = 4 * 85 + 1;
constexpr uint32_t kRingBit = 0x80000000u;

struct FrameRef {
    const uint8_t *host = nullptr;
    int stride = 0, pic = -1;
    std::vector<uint8_t> shadow;    // w*h bytes the device picture holds
    uint64_t last_use = 0;
};

// The search key of a call on the frame's planes: reference slot, block
// size, node position (multiples of 8), displacement of the centre from the
// node (|d| < 2048) and search range (< 16384).  Never 0.
bool frame_key(int r, int px, int py, int cx, int cy, int n, int sr, uint64_t &k)
{
    const int code = n == 8 ? 0 : n == 16 ? 1 : n == 32 ? 2 : n == 64 ? 3 : -1;
    if (code < 0 || r < 0 || r > 3 || px < 0 || py < 0 || (px & 7) || (py & 7) ||
        px >= tcme::kMaxPicSide || py >= tcme::kMaxPicSide || sr < 0 || sr > 16383)
        return false;
    const int64_t dx = static_cast<int64_t>(cx) - px + 2048, dy = static_cast<int64_t>(cy) - py + 2048;
    if (dx < 1 || dx > 4095 || dy < 1 || dy > 4095) return false;
    k = static_cast<uint64_t>(r) | static_cast<uint64_t>(code) << 2 |
        static_cast<uint64_t>(px >> 3) << 4 | static_cast<uint64_t>(py >> 3) << 15 |
        static_cast<uint64_t>(dx) << 26 | static_cast<uint64_t>(dy) << 38 |
        static_cast<uint64_t>(sr) << 50;
    return true;
}

uint32_t key_hash(uint64_t k)
{
    k ^= k >> 29;
    k *= 0xBF58476D1CE4E5B9ull;
    k ^= k >> 32;
    return static_cast<uint32_t>(k);
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

struct tcme_frame {
    std::mutex mu;
    int w = 0, h = 0;
    int cur_pic = -1;
    const uint8_t *cur_host = nullptr;
    int cur_stride = 0;
    std::vector<FrameRef> pool;     // device pictures of references, reused by content
    uint64_t tick = 0;
    int nrefs = 0;
    const uint8_t *ref_host[kFrameRefs] = {};
    int ref_pic[kFrameRefs] = {};
    std::vector<tcme_spec> specs;   // grids not kept
    int active = 0;                 // a frame is declared
    // frame-start searches: host list (pinned), device list and results,
    // results copied back to pinned memory
    tcme_call *h_calls = nullptr;
    tcme_result *h_res = nullptr;
    tcme_call *d_calls = nullptr;
    tcme_result *d_res = nullptr;
    size_t cap = 0;
    int nspec = 0;
    std::vector<uint64_t> spec_key;
    hipEvent_t spec_done = nullptr;
    bool spec_wait = false;
    // miss batches: calls and results in mapped pinned memory
    tcme_call *ring_calls = nullptr, *ring_calls_d = nullptr;
    tcme_result *ring_res = nullptr, *ring_res_d = nullptr;
    std::vector<uint64_t> ring_key;
    int ring_pos = 0;
    hipStream_t st[2] = {nullptr, nullptr};
    int next = 0;
    // open-addressing table: key -> result index (kRingBit: a ring slot)
    std::vector<uint64_t> hkey;
    std::vector<uint32_t> hval;
    uint32_t hmask = 0;
    size_t hcount = 0;
    tcme_frame_stats stats{};
};

namespace {

bool frame_find(const tcme_frame *f, uint64_t k, uint32_t &v)
{
    if (!f->hmask) return false;
    for (uint32_t i = key_hash(k) & f->hmask;; i = (i + 1) & f->hmask) {
        if (f->hkey[i] == k) { v = f->hval[i]; return true; }
        if (f->hkey[i] == 0) return false;
    }
}

// Inserts k -> v; a present k keeps its value unless `replace`.  Returns
// whether k was absent.
bool frame_insert(tcme_frame *f, uint64_t k, uint32_t v, bool replace = false)
{
    for (uint32_t i = key_hash(k) & f->hmask;; i = (i + 1) & f->hmask) {
        if (f->hkey[i] == k) {
            if (replace) f->hval[i] = v;
            return false;
        }
        if (f->hkey[i] == 0) {
            f->hkey[i] = k;
            f->hval[i] = v;
            f->hcount++;
            return true;
        }
    }
}

// A live result index for k: a frame-start search, or a ring slot that still
// holds k's call (a slot reused by a later batch no longer does).
bool frame_lookup(const tcme_frame *f, uint64_t k, uint32_t &v)
{
    return frame_find(f, k, v) && (!(v & kRingBit) || f->ring_key[v & ~kRingBit] == k);
}

void frame_table_reset(tcme_frame *f, size_t entries)
{
    size_t cap = 1024;
    while (cap < 2 * (entries + kRing)) cap <<= 1;
    if (cap != f->hkey.size()) {
        f->hkey.assign(cap, 0);
        f->hval.assign(cap, 0);
    } else {
        std::fill(f->hkey.begin(), f->hkey.end(), 0);
    }
    f->hmask = static_cast<uint32_t>(cap - 1);
    f->hcount = 0;
}

// Rebuilds the table from the frame-start searches and the live ring slots
// (stale ring keys are dropped).
void frame_table_rebuild(tcme_frame *f)
{
    std::fill(f->hkey.begin(), f->hkey.end(), 0);
    f->hcount = 0;
    for (int i = 0; i < f->nspec; ++i) frame_insert(f, f->spec_key[i], static_cast<uint32_t>(i));
    for (int i = 0; i < kRing; ++i)
        if (f->ring_key[i]) frame_insert(f, f->ring_key[i], kRingBit | i);
}

int frame_alloc_lists(tcme_frame *f, size_t n)
{
    if (n <= f->cap) return TCME_OK;
    HIP_TRY(hipDeviceSynchronize());
    if (f->h_calls) (void)hipHostFree(f->h_calls);
    if (f->h_res) (void)hipHostFree(f->h_res);
    (void)hipFree(f->d_calls);
    (void)hipFree(f->d_res);
    f->h_calls = nullptr; f->h_res = nullptr; f->d_calls = nullptr; f->d_res = nullptr;
    f->cap = 0;
    const size_t cap = std::max<size_t>(n, 4096);
    if (hipHostMalloc(reinterpret_cast<void **>(&f->h_calls), cap * sizeof(tcme_call), 0) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&f->h_res), cap * sizeof(tcme_result), 0) != hipSuccess ||
        hipMalloc(&f->d_calls, cap * sizeof(tcme_call)) != hipSuccess ||
        hipMalloc(&f->d_res, cap * sizeof(tcme_result)) != hipSuccess)
        return fail(TCME_E_NOMEM, "frame search lists for %zu calls", cap);
    f->cap = cap;
    return TCME_OK;
}

// The pool picture holding host plane `ref` (exactly its bytes), uploading
// it into the least recently used picture not in `busy` when none does.
int frame_ref_pic(tcme_frame *f, const uint8_t *ref, int stride, const int *busy, int nbusy,
                  hipStream_t s)
{
    tcme_ctx *c = g_drop.ctx;
    const size_t w = static_cast<size_t>(f->w);
    for (FrameRef &e : f->pool) {
        if (e.host != ref || e.stride != stride || e.pic < 0) continue;
        bool same = true;
        for (int y = 0; y < f->h && same; ++y)
            same = memcmp(e.shadow.data() + y * w, ref + static_cast<size_t>(y) * stride, w) == 0;
        if (same) {
            e.last_use = ++f->tick;
            return e.pic;
        }
    }
    FrameRef *pick = nullptr;
    for (FrameRef &e : f->pool) {
        bool used = false;
        for (int i = 0; i < nbusy; ++i) used |= e.pic == busy[i];
        if (!used && (!pick || e.last_use < pick->last_use)) pick = &e;
    }
    if (!pick) {
        f->pool.emplace_back();
        pick = &f->pool.back();
        pick->pic = pic_alloc(c, f->w, f->h, true);
        if (pick->pic < 0) return pick->pic;
    }
    pick->host = ref;
    pick->stride = stride;
    pick->shadow.resize(w * f->h);
    for (int y = 0; y < f->h; ++y)
        memcpy(pick->shadow.data() + y * w, ref + static_cast<size_t>(y) * stride, w);
    if (int e = tcme_pic_upload(c, pick->pic, pick->shadow.data(), f->w, s)) return e;
    pick->last_use = ++f->tick;
    return pick->pic;
}

void frame_release_pics(tcme_frame *f)
{
    tcme_ctx *c = g_drop.ctx;
    if (!c) return;
    for (FrameRef &e : f->pool)
        if (e.pic >= 0) pic_free(c, e.pic);
    f->pool.clear();
    if (f->cur_pic >= 0) pic_free(c, f->cur_pic);
    f->cur_pic = -1;
}

// Waits until ring slot i holds its result (status written last by the
// kernel); past 2 ms without it, synchronises the streams, which also
// reports a failed launch.
const tcme_result &ring_wait(tcme_frame *f, int i)
{
    volatile tcme_result *r = f->ring_res + i;
    const auto t0 = std::chrono::steady_clock::now();
    for (unsigned spin = 1; r->status == TCME_E_HIP; ++spin) {
        if ((spin & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::milliseconds(2)) {
            if (hipStreamSynchronize(f->st[0]) != hipSuccess ||
                hipStreamSynchronize(f->st[1]) != hipSuccess)
                die("frame search");
            if (r->status == TCME_E_HIP) {
                fail(TCME_E_HIP, "a frame search never wrote its result");
                die("frame search");
            }
            break;
        }
        __builtin_ia32_pause();
    }
    std::atomic_thread_fence(std::memory_order_acquire);
    return f->ring_res[i];
}

}  // namespace

TCME_API void tcme_prepare(void)
{
    static std::once_flag started;
    std::call_once(started, [] { std::thread([] { std::call_once(g_drop_once, dropin_init_body); }).detach(); });
}

TCME_API int tcme_frame_create(tcme_frame **out)
{
    if (!out) return fail(TCME_E_ARG, "null out");
    *out = nullptr;
    dropin_ensure();
    tcme_frame *f = new tcme_frame();
    if (hipSetDevice(g_drop.ctx->device) != hipSuccess ||
        hipEventCreateWithFlags(&f->spec_done, hipEventDisableTiming) != hipSuccess ||
        hipStreamCreateWithFlags(&f->st[0], hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&f->st[1], hipStreamNonBlocking) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&f->ring_calls), kRing * sizeof(tcme_call),
                      hipHostMallocCoherent) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void **>(&f->ring_res), kRing * sizeof(tcme_result),
                      hipHostMallocCoherent) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&f->ring_calls_d), f->ring_calls, 0) != hipSuccess ||
        hipHostGetDevicePointer(reinterpret_cast<void **>(&f->ring_res_d), f->ring_res, 0) != hipSuccess) {
        tcme_frame_destroy(f);
        return fail(TCME_E_HIP, "frame streams and miss-batch memory");
    }
    f->ring_key.assign(kRing, 0);
    *out = f;
    return TCME_OK;
}

TCME_API void tcme_frame_destroy(tcme_frame *f)
{
    if (!f) return;
    if (g_drop.ctx) {
        (void)hipSetDevice(g_drop.ctx->device);
        for (hipStream_t s : f->st)
            if (s) (void)hipStreamSynchronize(s);
        std::lock_guard<std::mutex> g(g_drop.ctx->mu);
        frame_release_pics(f);
    }
    for (hipStream_t s : f->st)
        if (s) (void)hipStreamDestroy(s);
    if (f->spec_done) (void)hipEventDestroy(f->spec_done);
    if (f->h_calls) (void)hipHostFree(f->h_calls);
    if (f->h_res) (void)hipHostFree(f->h_res);
    (void)hipFree(f->d_calls);
    (void)hipFree(f->d_res);
    if (f->ring_calls) (void)hipHostFree(f->ring_calls);
    if (f->ring_res) (void)hipHostFree(f->ring_res);
    delete f;
}

TCME_API int tcme_frame_begin(tcme_frame *f, const uint8_t *cur, int cur_stride, int w, int h,
                              const uint8_t *const *refs, const int *ref_strides, int nrefs,
                              const tcme_spec *specs, int nspecs)
{
    if (!f) return fail(TCME_E_ARG, "null frame");
    if (!cur || w <= 0 || h <= 0 || w > tcme::kMaxPicSide || h > tcme::kMaxPicSide || cur_stride < w)
        return fail(TCME_E_ARG, "bad current plane %dx%d stride %d", w, h, cur_stride);
    if (nrefs < 0 || nrefs > kFrameRefs || (nrefs > 0 && (!refs || !ref_strides)))
        return fail(TCME_E_ARG, "bad reference list (%d, at most %d)", nrefs, kFrameRefs);
    if (nspecs < 0 || (nspecs > 0 && !specs)) return fail(TCME_E_ARG, "bad spec list");
    for (int i = 0; i < nrefs; ++i)
        if (refs[i] && ref_strides[i] < w) return fail(TCME_E_ARG, "reference %d stride", i);
    for (int i = 0; i < nspecs; ++i) {
        const tcme_spec &p = specs[i];
        if (p.ref < 0 || p.ref >= nrefs || p.ref > 3 || !refs[p.ref] || p.sr < 0 || p.sr > 16383 ||
            (p.sizes & ~15) || (p.mode != TCME_CENTRE_OFFSET && p.mode != TCME_CENTRE_GRID) ||
            (p.mode == TCME_CENTRE_GRID && (!p.grid || p.grid_w < 0 || p.grid_h < 0)) ||
            p.dx < -8188 || p.dx > 8188 || p.dy < -8188 || p.dy > 8188)
            return fail(TCME_E_ARG, "bad spec %d", i);
    }
    const auto t0 = std::chrono::steady_clock::now();
    std::lock_guard<std::mutex> g(f->mu);
    tcme_ctx *c = g_drop.ctx;
    HIP_TRY(hipSetDevice(c->device));
    // the previous frame's work may still read the pictures
    HIP_TRY(hipStreamSynchronize(f->st[0]));
    HIP_TRY(hipStreamSynchronize(f->st[1]));
    std::lock_guard<std::mutex> gc(c->mu);
    if (w != f->w || h != f->h) {
        frame_release_pics(f);
        f->w = w;
        f->h = h;
    }
    if (f->cur_pic < 0) {
        f->cur_pic = pic_alloc(c, w, h, false);
        if (f->cur_pic < 0) return f->cur_pic;
    }
    hipStream_t s = f->st[0];
    if (int e = tcme_pic_upload(c, f->cur_pic, cur, cur_stride, s)) return e;
    f->cur_host = cur;
    f->cur_stride = cur_stride;
    f->nrefs = nrefs;
    int busy[kFrameRefs];
    int nbusy = 0;
    for (int i = 0; i < nrefs; ++i) {
        f->ref_host[i] = refs[i];
        f->ref_pic[i] = -1;
        if (!refs[i]) continue;
        const int pic = frame_ref_pic(f, refs[i], ref_strides[i], busy, nbusy, s);
        if (pic < 0) return pic;
        f->ref_pic[i] = pic;
        busy[nbusy++] = pic;
    }
    // the frame-start searches, deduplicated by key
    size_t upper = 0;
    for (int i = 0; i < nspecs; ++i)
        for (int k = 0; k < 4; ++k)
            if (specs[i].sizes >> k & 1) upper += static_cast<size_t>(w / (8 << k)) * (h / (8 << k));
    if (int e = frame_alloc_lists(f, upper + 1)) return e;
    frame_table_reset(f, upper);
    f->spec_key.resize(upper + 1);
    std::fill(f->ring_key.begin(), f->ring_key.end(), 0);
    f->ring_pos = 0;
    int n = 0;
    for (int i = 0; i < nspecs; ++i) {
        const tcme_spec &p = specs[i];
        for (int k = 0; k < 4; ++k) {
            if (!(p.sizes >> k & 1)) continue;
            const int bs = 8 << k;
            for (int py = 0; py + bs <= h; py += bs)
                for (int px = 0; px + bs <= w; px += bs) {
                    int dx = p.dx, dy = p.dy;
                    if (p.mode == TCME_CENTRE_GRID) {
                        const int gx = px >> 3, gy = py >> 3;
                        if (gx >= p.grid_w || gy >= p.grid_h) continue;
                        const tcme_cell &cell = p.grid[static_cast<size_t>(gy) * p.grid_w + gx];
                        if (cell.unavailable) continue;
                        dx = cell.dx;
                        dy = cell.dy;
                    }
                    const int cx = (4 * px + dx) >> 2, cy = (4 * py + dy) >> 2;
                    uint64_t key;
                    if (!frame_key(p.ref, px, py, cx, cy, bs, p.sr, key)) continue;
                    if (!frame_insert(f, key, static_cast<uint32_t>(n))) continue;
                    f->spec_key[n] = key;
                    f->h_calls[n++] = tcme_call{f->ref_pic[p.ref], f->cur_pic, px, py, cx, cy, bs, p.sr};
                }
        }
    }
    f->nspec = n;
    f->specs.assign(specs, specs + nspecs);
    if (n > 0) {
        HIP_TRY(hipMemcpyAsync(f->d_calls, f->h_calls, n * sizeof(tcme_call), hipMemcpyHostToDevice, s));
        if (int e = dispatch(c, f->d_calls, n, c->tiling, f->d_res, nullptr, s)) return e;
        HIP_TRY(hipMemcpyAsync(f->h_res, f->d_res, n * sizeof(tcme_result), hipMemcpyDeviceToHost, s));
    }
    HIP_TRY(hipEventRecord(f->spec_done, s));
    f->spec_wait = true;
    f->active = 1;
    f->stats.frames++;
    f->stats.spec_calls += n;
    f->stats.begin_s += seconds_since(t0);
    return TCME_OK;
}

namespace {

// One frame call that found no result: searched now, with the other nodes of
// its CTU against the same references at the same displacement.
tcme_result frame_miss(tcme_frame *f, int r, int px, int py, int cx, int cy, int n, int sr)
{
    const auto t0 = std::chrono::steady_clock::now();
    // the references the frame searches at the predictor centre with this range
    int refs[kFrameRefs];
    int nr = 0;
    refs[nr++] = r;
    for (const tcme_spec &p : f->specs)
        if (p.mode == TCME_CENTRE_OFFSET && p.dx == 0 && p.dy == 0 && p.sr == sr &&
            std::find(refs, refs + nr, p.ref) == refs + nr)
            refs[nr++] = p.ref;
    if (f->ring_pos + kBatchMax > kRing) {
        // wrap: every earlier batch must be done before its slots are reused
        if (hipStreamSynchronize(f->st[0]) != hipSuccess || hipStreamSynchronize(f->st[1]) != hipSuccess)
            die("frame search");
        f->ring_pos = 0;
    }
    const int base = f->ring_pos;
    int k = 0;
    auto add = [&](int rr, int x, int y, int bs, bool first) {
        uint64_t key;
        if (!frame_key(rr, x, y, x + (cx - px), y + (cy - py), bs, sr, key)) return;
        uint32_t v;
        if (!first && frame_lookup(f, key, v)) return;
        const int i = base + k++;
        f->ring_calls[i] = tcme_call{f->ref_pic[rr], f->cur_pic, x, y, x + (cx - px), y + (cy - py), bs, sr};
        f->ring_res[i].status = TCME_E_HIP;
        // the slot's previous call (if any) is no longer found: lookups check
        // that the slot still holds the key they look for
        f->ring_key[i] = key;
        if (f->hcount * 4 > f->hkey.size() * 3) frame_table_rebuild(f);
        frame_insert(f, key, kRingBit | static_cast<uint32_t>(i), true);
    };
    add(r, px, py, n, true);
    const int x0 = px & ~63, y0 = py & ~63;
    for (int q = 0; q < nr; ++q)
        for (int bs = 64; bs >= 8; bs >>= 1)
            for (int y = y0; y < y0 + 64 && y + bs <= f->h; y += bs)
                for (int x = x0; x < x0 + 64 && x + bs <= f->w; x += bs)
                    if (!(refs[q] == r && x == px && y == py && bs == n) && f->ref_pic[refs[q]] >= 0)
                        add(refs[q], x, y, bs, false);
    f->ring_pos += k;
    hipStream_t s = f->st[f->next];
    f->next ^= 1;
    if (tcme::launch_list(g_drop.ctx->d_table, f->ring_calls_d + base, k, f->ring_res_d + base, s) !=
        hipSuccess)
        die("frame search launch");
    const tcme_result res = ring_wait(f, base);
    f->stats.misses++;
    f->stats.batch_calls += k;
    f->stats.miss_s += seconds_since(t0);
    return res;
}

}  // namespace

TCME_API tcme_mv tcme_frame_motion_est(tcme_frame *f, const uint8_t *ref, int ref_stride, int ref_w,
                                       int ref_h, const uint8_t *orig, int orig_stride,
                                       int center_x, int center_y, int blk_size,
                                       int search_range, int32_t *best_sad_out)
{
    if (!f) {
        fail(TCME_E_ARG, "null frame");
        die("tcme_frame_motion_est");
    }
    std::unique_lock<std::mutex> g(f->mu);
    f->stats.calls++;
    // a call on the declared planes: which reference, and where the block is
    int r = -1;
    if (f->active && ref_w == f->w && ref_h == f->h && orig_stride == f->cur_stride)
        for (int i = 0; i < f->nrefs && r < 0; ++i)
            if (f->ref_host[i] == ref && f->ref_pic[i] >= 0) r = i;
    int px = -1, py = -1;
    if (r >= 0 && orig >= f->cur_host &&
        orig < f->cur_host + static_cast<size_t>(f->cur_stride) * f->h) {
        const size_t off = static_cast<size_t>(orig - f->cur_host);
        px = static_cast<int>(off % f->cur_stride);
        py = static_cast<int>(off / f->cur_stride);
    }
    uint64_t key;
    if (r < 0 || px < 0 || px > f->w - blk_size || py > f->h - blk_size ||
        !frame_key(r, px, py, center_x, center_y, blk_size, search_range, key)) {
        // not a node of the declared frame: the per-call drop-in
        f->stats.uncached++;
        g.unlock();
        return tc_motion_est(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x, center_y,
                             blk_size, search_range, best_sad_out);
    }
    if (f->spec_wait) {
        const auto t0 = std::chrono::steady_clock::now();
        if (hipEventSynchronize(f->spec_done) != hipSuccess) die("frame-start search");
        f->spec_wait = false;
        f->stats.wait_s += seconds_since(t0);
    }
    tcme_result res;
    uint32_t v;
    if (frame_lookup(f, key, v) && !(v & kRingBit)) {
        res = f->h_res[v];
        f->stats.hits_spec++;
    } else if (frame_lookup(f, key, v)) {
        res = ring_wait(f, static_cast<int>(v & ~kRingBit));
        f->stats.hits_batch++;
    } else {
        res = frame_miss(f, r, px, py, center_x, center_y, blk_size, search_range);
    }
    if (res.status != 0) {
        fail(res.status, "frame search rejected a call (status %d)", res.status);
        die("tcme_frame_motion_est");
    }
    if (best_sad_out) *best_sad_out = res.sad;
    return tcme_mv{res.mv_x, res.mv_y};
}

TCME_API int tcme_frame_end(tcme_frame *f)
{
    if (!f) return fail(TCME_E_ARG, "null frame");
    std::lock_guard<std::mutex> g(f->mu);
    f->active = 0;
    return TCME_OK;
}

TCME_API int tcme_frame_get_stats(tcme_frame *f, tcme_frame_stats *st)
{
    if (!f || !st) return fail(TCME_E_ARG, "null argument");
    std::lock_guard<std::mutex> g(f->mu);
    *st = f->stats;
    return TCME_OK;
}
