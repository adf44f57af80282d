/*
 * c_abi_search.c - the batch API from plain C (no HIP or Python on the caller's
 * side): uploads a reference and a current luma plane, searches a list of calls
 * with tcme_search_host, and writes the results.  Built and run by
 * tests/test_c_abi.py.
 *
 * usage: c_abi_search W H ref.y cur.y calls.i32 n out.i32
 *   ref.y, cur.y  W*H bytes each
 *   calls.i32     n rows of int32 (px, py, cx, cy, blk, sr)
 *   out.i32       written: n rows of int32 (mv_x, mv_y, sad, status)
 */
#include <stdio.h>
#include <stdlib.h>

#include "tcme.h"

static void *slurp(const char *path, size_t bytes)
{
    FILE *f = fopen(path, "rb");
    void *p = malloc(bytes ? bytes : 1);
    if (!f || !p || fread(p, 1, bytes, f) != bytes) {
        fprintf(stderr, "cannot read %zu bytes of %s\n", bytes, path);
        exit(2);
    }
    fclose(f);
    return p;
}

int main(int argc, char **argv)
{
    if (argc != 8) {
        fprintf(stderr, "usage: %s W H ref cur calls n out\n", argv[0]);
        return 2;
    }
    const int w = atoi(argv[1]), h = atoi(argv[2]), n = atoi(argv[6]);
    uint8_t *ref = slurp(argv[3], (size_t)w * h);
    uint8_t *cur = slurp(argv[4], (size_t)w * h);
    int32_t *rows = slurp(argv[5], (size_t)n * 6 * sizeof(int32_t));

    tcme_ctx *ctx = NULL;
    if (tcme_create(&ctx, -1, 16) != TCME_OK) {
        fprintf(stderr, "tcme_create: %s\n", tcme_last_error());
        return 1;
    }
    const int rp = tcme_pic_create(ctx, w, h);
    const int cp = tcme_pic_create_plane(ctx, w, h);
    if (rp < 0 || cp < 0 || tcme_pic_upload(ctx, rp, ref, w, NULL) != TCME_OK ||
        tcme_pic_upload(ctx, cp, cur, w, NULL) != TCME_OK) {
        fprintf(stderr, "pictures: %s\n", tcme_last_error());
        return 1;
    }
    tcme_call *calls = malloc((size_t)n * sizeof *calls);
    tcme_result *res = malloc((size_t)n * sizeof *res);
    for (int i = 0; i < n; i++) {
        const int32_t *r = rows + 6 * i;
        calls[i].ref_pic = rp;
        calls[i].cur_pic = cp;
        calls[i].px = r[0];
        calls[i].py = r[1];
        calls[i].cx = r[2];
        calls[i].cy = r[3];
        calls[i].blk = r[4];
        calls[i].sr = r[5];
    }
    if (tcme_search_host(ctx, calls, n, res, NULL, NULL) != TCME_OK) {
        fprintf(stderr, "tcme_search_host: %s\n", tcme_last_error());
        return 1;
    }
    FILE *f = fopen(argv[7], "wb");
    if (!f || fwrite(res, sizeof *res, (size_t)n, f) != (size_t)n) {
        fprintf(stderr, "cannot write %s\n", argv[7]);
        return 2;
    }
    fclose(f);
    tcme_destroy(ctx);
    printf("%s: %d calls searched\n", tcme_version(), n);
    return 0;
}
