/*
 * me_shim.c - checker harness linked into oracle/_ref/tcenc_hook.
 *
 * The reference encoder is compiled unchanged except that, on the compiler
 * command line, motion.c's tc_motion_est is renamed tc_motion_est_ref and
 * tcenc.c's tc_encoder_create is renamed tcme_shim_encoder_create.  This file
 * then supplies both symbols.  The ME call site contract it sits on is
 * src/motion.c:384-389 / include/tcodec_common.h:454-459.
 *
 * Built with -DTCME_SHIM_FRAME (oracle/_ref/tcenc_fcheck) the file instead
 * supplies tcme_shim_motion_est, the name encoder.c's calls get on its
 * compile line there, and the GPU side of every call is the frame-level hook
 * of tcodec_amd/glue/tcme_encoder_hook.c (linked with libtcme.so) instead of
 * the library named by TCME_LIB.
 *
 * TCME_SHIM selects what each tc_motion_est call does:
 *   ref    (default) run the reference function, time it (CLOCK_MONOTONIC)
 *   gpu    run tc_motion_est from the library named by TCME_LIB (libtcme.so)
 *   check  run both, compare {x, y, sad}, return the reference answer
 *   trace  run the reference and append every call (inputs, outputs and the
 *          luma planes it reads, keyed by POC) to the file named by TCME_TRACE
 * At exit one line goes to stderr:
 *   TCME_SHIM mode=<m> calls=<n> mismatches=<k> me_seconds=<t> first_call_seconds=<f>
 * (me_seconds includes the first call, which in gpu mode also starts the GPU
 * runtime), then one line per call site (the caller's return address; a site
 * the compiler duplicated by inlining counts once per copy):
 *   TCME_SITE ra=<offset> calls=<n> slot1=<a> glob_centre=<g> bframe=<b>
 * slot1: calls against dpb[1] (the multi-reference search, encoder.c:982);
 * glob_centre: calls centred on the global-motion centre of a P frame with
 * |glob| > 16 quarter samples (encoder.c:1000-1004); bframe: calls against a
 * picture that is not the most recent reference (B forward/backward,
 * encoder.c:807, 834, or dpb[1..3]).  Counters are atomic: the v1 encoder
 * calls from its WPP pool threads (encoder.c:3014).
 */
#include "tcodec_common.h"
#include <dlfcn.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

tc_mv_s tc_motion_est_ref(const tc_pixel_t *ref, int ref_stride, int ref_w, int ref_h,
                          const tc_pixel_t *orig, int orig_stride, int center_x,
                          int center_y, int blk_size, int search_range,
                          tc_sad_t *best_sad_out);
tc_encoder_t *tc_encoder_create(const tc_config_t *config);

typedef tc_mv_s (*me_fn_t)(const tc_pixel_t *, int, int, int, const tc_pixel_t *, int,
                           int, int, int, int, tc_sad_t *);

#ifdef TCME_SHIM_FRAME
#define SHIM_ENTRY tcme_shim_motion_est
tc_mv_s tcme_hook_motion_est(const tc_pixel_t *ref, int ref_stride, int ref_w, int ref_h,
                             const tc_pixel_t *orig, int orig_stride, int center_x,
                             int center_y, int blk_size, int search_range,
                             tc_sad_t *best_sad_out);
tc_mv_s SHIM_ENTRY(const tc_pixel_t *ref, int ref_stride, int ref_w, int ref_h,
                   const tc_pixel_t *orig, int orig_stride, int center_x, int center_y,
                   int blk_size, int search_range, tc_sad_t *best_sad_out);
#else
#define SHIM_ENTRY tc_motion_est
#endif

enum { MODE_REF, MODE_GPU, MODE_CHECK, MODE_TRACE };
static int g_mode = -1;
static me_fn_t g_gpu_fn;
static long g_calls, g_mism;
static double g_sec, g_first;
static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;  /* sites, timing, trace */
#define MAX_SITES 32
static struct { void *ra; long calls, slot1, glob, bframe; } g_site[MAX_SITES];
static int g_nsites;
static tc_encoder_t *g_enc;
static FILE *g_trace;
#define MAX_SEEN 4096
static int g_seen_ref[MAX_SEEN];
static int g_cur_id = -1;
static unsigned char *g_cur_shadow;

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static const char *mode_name(void)
{
    static const char *names[] = {"ref", "gpu", "check", "trace"};
    return g_mode >= 0 ? names[g_mode] : "unset";
}

static void shim_report(void)
{
    fprintf(stderr, "TCME_SHIM mode=%s calls=%ld mismatches=%ld me_seconds=%.6f "
            "first_call_seconds=%.6f\n", mode_name(), g_calls, g_mism, g_sec, g_first);
    Dl_info di;
    for (int i = 0; i < g_nsites; i++) {
        long off = (dladdr(g_site[i].ra, &di) && di.dli_fbase)
                       ? (long)((const char *)g_site[i].ra - (const char *)di.dli_fbase) : 0;
        fprintf(stderr, "TCME_SITE ra=0x%lx calls=%ld slot1=%ld glob_centre=%ld bframe=%ld\n",
                off, g_site[i].calls, g_site[i].slot1, g_site[i].glob, g_site[i].bframe);
    }
    if (g_trace) fclose(g_trace);
}

/* Books a call against the site that made it (see the header comment). */
static void count_site(void *ra, const tc_pixel_t *ref, const tc_pixel_t *orig, int cx, int cy)
{
    int slot = -1, glob = 0;
    if (g_enc) {
        for (int i = 0; i < TC_REF_FRAMES; i++)
            if (g_enc->dpb[i].frame && g_enc->dpb[i].frame->y == ref) slot = i;
        const tc_frame_buf_t *cur = g_enc->cur;
        if (slot == 0 && cur && g_enc->glob_mv_valid &&
            (g_enc->glob_mv_x > 16 || g_enc->glob_mv_x < -16 ||
             g_enc->glob_mv_y > 16 || g_enc->glob_mv_y < -16) &&
            orig >= cur->y && orig < cur->y + (size_t)cur->stride_y * cur->height) {
            size_t off = (size_t)(orig - cur->y);
            int px = (int)(off % (size_t)cur->stride_y), py = (int)(off / (size_t)cur->stride_y);
            glob = cx == ((px * 4 + g_enc->glob_mv_x) >> 2) &&
                   cy == ((py * 4 + g_enc->glob_mv_y) >> 2);
        }
    }
    pthread_mutex_lock(&g_mu);
    int k = 0;
    while (k < g_nsites && g_site[k].ra != ra) k++;
    if (k == g_nsites && g_nsites < MAX_SITES) g_site[g_nsites++].ra = ra;
    if (k < MAX_SITES) {
        g_site[k].calls++;
        g_site[k].slot1 += slot == 1;
        g_site[k].glob += glob;
        g_site[k].bframe += slot != 0;
    }
    pthread_mutex_unlock(&g_mu);
}

static void shim_init(void)
{
    const char *m = getenv("TCME_SHIM");
    g_mode = MODE_REF;
    if (m && !strcmp(m, "gpu")) g_mode = MODE_GPU;
    else if (m && !strcmp(m, "check")) g_mode = MODE_CHECK;
    else if (m && !strcmp(m, "trace")) g_mode = MODE_TRACE;
#ifdef TCME_SHIM_FRAME
    if (g_mode == MODE_GPU || g_mode == MODE_CHECK) g_gpu_fn = tcme_hook_motion_est;
#endif
    if (!g_gpu_fn && (g_mode == MODE_GPU || g_mode == MODE_CHECK)) {
        const char *lib = getenv("TCME_LIB");
        void *h = lib ? dlopen(lib, RTLD_NOW | RTLD_LOCAL) : NULL;
        g_gpu_fn = h ? (me_fn_t)dlsym(h, "tc_motion_est") : NULL;
        if (!g_gpu_fn) {
            fprintf(stderr, "TCME_SHIM: cannot load tc_motion_est from TCME_LIB=%s: %s\n",
                    lib ? lib : "(unset)", dlerror());
            exit(3);
        }
    }
    if (g_mode == MODE_TRACE) {
        const char *path = getenv("TCME_TRACE");
        g_trace = path ? fopen(path, "wb") : NULL;
        if (!g_trace) {
            fprintf(stderr, "TCME_SHIM: trace mode needs a writable TCME_TRACE\n");
            exit(3);
        }
    }
    atexit(shim_report);
}

tc_encoder_t *tcme_shim_encoder_create(const tc_config_t *config)
{
    g_enc = tc_encoder_create(config);
    return g_enc;
}

/* Trace records: a 4-int header {tag, a, b, c} followed by a payload.
 *   tag 1: plane  {1, kind(0 recon ref, 1 source cur), id, w*h} + w*h bytes (rows of w);
 *                 id is the POC for refs and the coding-order index for sources
 *   tag 2: call   {2, 16, 0, 0} + 16 int32:
 *          cur_id px py ref_poc cx cy blk sr w h mv_x mv_y sad orig_stride 0 0   */
static void trace_plane(int kind, int poc, const tc_pixel_t *p, int stride, int w, int h)
{
    int hdr[4] = {1, kind, poc, w * h};
    fwrite(hdr, sizeof(int), 4, g_trace);
    for (int y = 0; y < h; y++) fwrite(p + (size_t)y * stride, 1, (size_t)w, g_trace);
}

static void trace_call(const tc_pixel_t *ref, int ref_stride, int w, int h,
                       const tc_pixel_t *orig, int orig_stride, int cx, int cy,
                       int blk, int sr, tc_mv_s mv, tc_sad_t sad)
{
    int ref_poc = -1, cur_poc = -1, px = -1, py = -1;
    if (g_enc) {
        for (int i = 0; i < TC_REF_FRAMES; i++)
            if (g_enc->dpb[i].frame && g_enc->dpb[i].frame->y == ref) ref_poc = g_enc->dpb[i].poc;
        /* enc->cur is one buffer refilled per coded frame (encoder.c:2632),
         * so source planes are numbered by content change, in coding order. */
        const tc_frame_buf_t *cur = g_enc->cur;
        if (cur && orig >= cur->y && orig < cur->y + (size_t)cur->stride_y * cur->height) {
            size_t off = (size_t)(orig - cur->y);
            px = (int)(off % (size_t)cur->stride_y);
            py = (int)(off / (size_t)cur->stride_y);
            size_t plane = (size_t)w * h;
            int changed = (g_cur_id < 0);
            if (!g_cur_shadow) g_cur_shadow = (unsigned char *)malloc(plane);
            for (int y = 0; y < h && !changed; y++)
                changed = memcmp(g_cur_shadow + (size_t)y * w, cur->y + (size_t)y * cur->stride_y, (size_t)w) != 0;
            if (changed) {
                for (int y = 0; y < h; y++)
                    memcpy(g_cur_shadow + (size_t)y * w, cur->y + (size_t)y * cur->stride_y, (size_t)w);
                g_cur_id++;
                trace_plane(1, g_cur_id, cur->y, cur->stride_y, w, h);
            }
            cur_poc = g_cur_id;
        }
    }
    if (ref_poc >= 0 && ref_poc < MAX_SEEN && !g_seen_ref[ref_poc]) {
        g_seen_ref[ref_poc] = 1;
        trace_plane(0, ref_poc, ref, ref_stride, w, h);
    }
    int hdr[4] = {2, 16, 0, 0};
    int rec[16] = {cur_poc, px, py, ref_poc, cx, cy, blk, sr, w, h,
                   mv.x, mv.y, sad, orig_stride, 0, 0};
    fwrite(hdr, sizeof(int), 4, g_trace);
    fwrite(rec, sizeof(int), 16, g_trace);
}

tc_mv_s SHIM_ENTRY(const tc_pixel_t *ref, int ref_stride, int ref_w, int ref_h,
                   const tc_pixel_t *orig, int orig_stride, int center_x,
                   int center_y, int blk_size, int search_range,
                   tc_sad_t *best_sad_out)
{
    pthread_once(&g_once, shim_init);
    const long call = __atomic_add_fetch(&g_calls, 1, __ATOMIC_RELAXED);
    count_site(__builtin_return_address(0), ref, orig, center_x, center_y);
    tc_sad_t sad = 0;
    tc_mv_s mv;
    double t0 = now_s();
    if (g_mode == MODE_GPU) {
        mv = g_gpu_fn(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x, center_y,
                      blk_size, search_range, &sad);
    } else {
        mv = tc_motion_est_ref(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x,
                               center_y, blk_size, search_range, &sad);
    }
    const double dt = now_s() - t0;
    pthread_mutex_lock(&g_mu);
    g_sec += dt;
    if (call == 1) g_first = dt;
    pthread_mutex_unlock(&g_mu);
    if (g_mode == MODE_CHECK) {
        tc_sad_t gsad = 0;
        tc_mv_s gmv = g_gpu_fn(ref, ref_stride, ref_w, ref_h, orig, orig_stride,
                               center_x, center_y, blk_size, search_range, &gsad);
        if (gmv.x != mv.x || gmv.y != mv.y || gsad != sad) {
            if (__atomic_add_fetch(&g_mism, 1, __ATOMIC_RELAXED) <= 10)
                fprintf(stderr, "TCME_SHIM mismatch call=%ld c=(%d,%d) n=%d sr=%d "
                        "ref=(%d,%d,%d) gpu=(%d,%d,%d)\n", call, center_x, center_y,
                        blk_size, search_range, mv.x, mv.y, sad, gmv.x, gmv.y, gsad);
        }
    } else if (g_mode == MODE_TRACE) {
        pthread_mutex_lock(&g_mu);
        trace_call(ref, ref_stride, ref_w, ref_h, orig, orig_stride, center_x, center_y,
                   blk_size, search_range, mv, sad);
        pthread_mutex_unlock(&g_mu);
    }
    if (best_sad_out) *best_sad_out = sad;
    return mv;
}
