/* Device-free entry points of libmdzg under the host sanitizers.
 *
 * tests/test_host_sanitize_cpu.py compiles the library's host C sources with
 * AddressSanitizer and UndefinedBehaviorSanitizer (gcc, host code only; the
 * device objects are linked as built) together with this driver, and runs
 * it on the CPU: every call must return a result or an error code, with no
 * sanitizer report and no leak.
 *
 *   rects     mdzg_rect_from_center / _coords over precisions 2..4096, sizes
 *             down to 1e-290, odd frame shapes, bad strings, short buffers
 *   scenes    every file named on the command line, then every prefix of it
 *             (truncated at each of up to 64 points), lines made very long,
 *             values replaced by garbage: load, palette, dims, rect, free
 *   palettes  the same for .map files given after "--maps"
 *   png       small images and bad arguments
 *   context   mdzg_create without a device (an error, cleanly) or with one
 *             (created and destroyed)
 *
 * usage: host_check <tmpdir> <scene.mdz>... --maps <palette.map>...
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mdzg.h>

#define CAP 4096
static char A[CAP], B[CAP], C[CAP], D[CAP];
static long calls, errors;

static void note(int rc)
{
    ++calls;
    if (rc < 0) {
        ++errors;
        if (!mdzg_strerror(rc))
            abort();
    }
}

static void rects(void)
{
    static const long precs[] = { 2, 3, 53, 63, 64, 65, 80, 127, 128, 129, 176, 184, 320,
                                  1000, 1024, 1025, 4096 };
    static const char* sizes[] = { "3.0", "1e-25", "1e-100", "1e-290", "4", "0.5e-7" };
    static const int dims[][2] = { { 1, 1 }, { 240, 180 }, { 3, 7 }, { 1920, 1080 },
                                   { 7, 3 }, { 3840, 2160 } };
    for (size_t p = 0; p < sizeof precs / sizeof *precs; ++p)
        for (size_t s = 0; s < sizeof sizes / sizeof *sizes; ++s)
            for (size_t d = 0; d < sizeof dims / sizeof *dims; ++d) {
                note(mdzg_rect_from_center("-0.743643887037158704752191506114774",
                                           "0.131825904205311970493132056385139", sizes[s],
                                           precs[p], dims[d][0], dims[d][1], A, B, C, D, CAP));
                note(mdzg_rect_from_center_coords("-0.75", "0", sizes[s], precs[p], dims[d][0],
                                                  dims[d][1], A, B, C, D, CAP));
            }
    /* bad input: each must come back as an error code */
    note(mdzg_rect_from_center("abc", "0", "1", 128, 10, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("", "", "", 128, 10, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1e999999999", 128, 10, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1", 128, 0, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1", 128, 10, -1, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1", 0, 10, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1", 1024, 10, 10, A, B, C, D, 8));
    note(mdzg_rect_from_center(NULL, "0", "1", 128, 10, 10, A, B, C, D, CAP));
    note(mdzg_rect_from_center("0", "0", "1", 128, 10, 10, NULL, B, C, D, CAP));
}

static char* slurp(const char* path, long* n)
{
    FILE* f = fopen(path, "rb");
    if (!f)
        return NULL;
    fseek(f, 0, SEEK_END);
    *n = ftell(f);
    fseek(f, 0, SEEK_SET);
    char* buf = malloc((size_t)*n + 1);
    if (buf && fread(buf, 1, (size_t)*n, f) != (size_t)*n) {
        free(buf);
        buf = NULL;
    }
    fclose(f);
    return buf;
}

static int spill(const char* path, const char* data, long n)
{
    FILE* f = fopen(path, "wb");
    if (!f)
        return -1;
    const int ok = fwrite(data, 1, (size_t)n, f) == (size_t)n;
    return (fclose(f) == 0 && ok) ? 0 : -1;
}

static void one_scene(const char* path)
{
    mdzg_scene sc;
    const int rc = mdzg_scene_load(path, &sc);
    note(rc);
    if (rc < 0)
        return;
    mdzg_palette pal;
    memset(&pal, 0, sizeof pal);
    note(mdzg_scene_palette(&sc, &pal));
    int w = 0, h = 0;
    note(mdzg_scene_dims(&sc, &w, &h));
    if (w > 0 && h > 0 && w < 100000 && h < 100000) {
        note(mdzg_scene_rect(&sc, w, h, 0, A, B, C, D, CAP));
        note(mdzg_scene_rect(&sc, w, h, MDZG_SCENE_FIT_RECT, A, B, C, D, CAP));
    }
    note(mdzg_scene_rect(&sc, 640, 480, 0, A, B, C, D, CAP));
    mdzg_scene_free(&sc);
}

static void one_map(const char* path)
{
    mdzg_palette pal;
    memset(&pal, 0, sizeof pal);
    note(mdzg_palette_load(path, &pal));
}

/* the file, then damaged copies of it in tmpdir */
static void mangle(const char* tmpdir, const char* path, void (*use)(const char*))
{
    use(path);
    long n = 0;
    char* src = slurp(path, &n);
    if (!src)
        return;
    char tmp[4096];
    snprintf(tmp, sizeof tmp, "%s/mangled", tmpdir);
    /* prefixes */
    const long step = n / 64 > 0 ? n / 64 : 1;
    for (long k = 0; k < n; k += step)
        if (spill(tmp, src, k) == 0)
            use(tmp);
    /* a line made very long, and digits turned into letters */
    char* big = malloc((size_t)n + 70000);
    if (big) {
        const long cut = n / 2;
        memcpy(big, src, (size_t)cut);
        memset(big + cut, '7', 65536);
        memcpy(big + cut + 65536, src + cut, (size_t)(n - cut));
        if (spill(tmp, big, n + 65536) == 0)
            use(tmp);
        memcpy(big, src, (size_t)n);
        for (long k = 0; k < n; ++k)
            if (big[k] >= '0' && big[k] <= '9' && (k % 3) == 0)
                big[k] = 'x';
        if (spill(tmp, big, n) == 0)
            use(tmp);
        free(big);
    }
    free(src);
    remove(tmp);
}

static void png(const char* tmpdir)
{
    char path[4096];
    snprintf(path, sizeof path, "%s/check.png", tmpdir);
    static const int dims[][2] = { { 1, 1 }, { 7, 3 }, { 64, 64 }, { 3, 257 } };
    for (size_t d = 0; d < sizeof dims / sizeof *dims; ++d) {
        const int w = dims[d][0], h = dims[d][1];
        uint32_t* rgb = malloc(sizeof(uint32_t) * (size_t)w * (size_t)h);
        if (!rgb)
            abort();
        for (int i = 0; i < w * h; ++i)
            rgb[i] = (uint32_t)i * 2654435761u;
        note(mdzg_png_write(path, rgb, w, h));
        free(rgb);
    }
    uint32_t one = 0;
    note(mdzg_png_write(path, &one, 0, 1));
    note(mdzg_png_write(path, NULL, 1, 1));
    note(mdzg_png_write("/nonexistent-dir/x.png", &one, 1, 1));
    remove(path);
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        fprintf(stderr, "usage: host_check <tmpdir> <scene.mdz>... --maps <palette.map>...\n");
        return 2;
    }
    const char* tmpdir = argv[1];
    if (mdzg_abi_version() <= 0)
        return 3;
    for (int code = -20; code <= 0; ++code)
        if (!mdzg_strerror(code))
            return 4;
    for (int path = -1; path <= 4; ++path)
        for (long p = -1; p <= 1100; p += 37)
            (void)mdzg_kernel_id(path, p);
    rects();
    int maps = 0;
    for (int i = 2; i < argc; ++i) {
        if (!strcmp(argv[i], "--maps")) {
            maps = 1;
            continue;
        }
        mangle(tmpdir, argv[i], maps ? one_map : one_scene);
    }
    one_scene("/nonexistent.mdz");
    one_map("/nonexistent.map");
    png(tmpdir);
    mdzg_ctx* ctx = NULL;
    const int rc = mdzg_create(0, &ctx);
    if (rc == MDZG_OK)
        mdzg_destroy(ctx);
    else if (ctx != NULL)
        return 5;
    printf("host_check calls %ld errors %ld create %d\n", calls, errors, rc);
    return 0;
}
