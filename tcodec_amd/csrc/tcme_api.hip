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
#include <mutex>
#include <string>
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
    std::mutex mu;
    std::mutex launch_mu;
};

namespace {

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
    long calls = 0, uploads = 0;
};

DropIn g_drop;
constexpr int kMirrors = 8;

[[noreturn]] void die(const char *what)
{
    fprintf(stderr, "tcme: fatal: %s: %s\n", what, tcme_last_error());
    abort();
}

void dropin_init()
{
    DropIn &d = g_drop;
    const char *dev = getenv("TCME_DEVICE");
    if (tcme_create(&d.ctx, dev ? atoi(dev) : 0, 64) != TCME_OK) die("no usable GPU");
    if (hipStreamCreateWithFlags(&d.stream, hipStreamNonBlocking) != hipSuccess) die("stream");
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
        hipHostGetDevicePointer(&dr, d.h_res, 0) != hipSuccess)
        die("drop-in staging");
    d.d_blk = static_cast<const uint8_t *>(db);
    d.d_res = static_cast<tcme_result *>(dr);
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
    std::lock_guard<std::mutex> g(d.mu);
    if (!d.ctx) dropin_init();
    if (blk_size != 8 && blk_size != 16 && blk_size != 32 && blk_size != 64) {
        fail(TCME_E_BLK, "block size %d (supported: 8, 16, 32, 64)", blk_size);
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
