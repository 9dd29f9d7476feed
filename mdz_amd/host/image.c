/*
 * image.c -- the image end of the command-line path: a whole frame rendered
 * and coloured in one device job (what render_to_file, src/render.c:10-105,
 * does with the line pool, palette_apply and do_anti_aliasing), and the PNG
 * file save_png_file writes (src/my_png.c:28-109).
 *
 * The PNG is 8-bit RGB, non-interlaced, from the low three bytes of each
 * rgb_data colour, as MDZ's png_set_filler(0, PNG_FILLER_AFTER) drops the
 * fourth.  Rows are filtered with the filter whose output has the smallest
 * sum of absolute byte values (the heuristic libpng applies by default to
 * 8-bit RGB) and deflated by zlib at its default level, so the file decodes
 * to the same pixels libpng's would, at a similar size.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <zlib.h>

#include "mdzg_host.h"

int mdzg_render_image(mdzg_ctx* ctx, const mdzg_view* view, int aa, const mdzg_palette* pal,
                      uint32_t* rgb, uint32_t* counts, mdzg_stats* stats)
{
    if (!ctx || !view || !rgb || aa < 1 || aa > 64 || !mdzg_palette_ok(pal) ||
        view->real_width < aa || view->real_height < aa || view->real_width % aa ||
        view->real_height % aa)
        return MDZG_ERR_ARG;
    mdzg_tile t;
    memset(&t, 0, sizeof t);
    t.view = *view;
    t.line0 = 0;
    t.nlines = view->real_height;
    t.col0 = 0;
    t.ncols = view->real_width;
    t.counts = counts ? counts : rgb;     /* the image path never scatters into it */
    mdzg_image_req r;
    r.user_w = view->real_width / aa;
    r.user_h = view->real_height / aa;
    r.aa = aa;
    r.pal = pal;
    r.rgb = rgb;
    r.counts = counts;
    return mdzg_rt_render_image(ctx, &t, &r, stats);
}

/* ---- PNG ------------------------------------------------------------------ */

static void put32(unsigned char* p, uint32_t v)
{
    p[0] = (unsigned char)(v >> 24);
    p[1] = (unsigned char)(v >> 16);
    p[2] = (unsigned char)(v >> 8);
    p[3] = (unsigned char)v;
}

static int chunk(FILE* f, const char* type, const unsigned char* data, uint32_t len)
{
    unsigned char hdr[8];
    put32(hdr, len);
    memcpy(hdr + 4, type, 4);
    uLong crc = crc32(0L, (const Bytef*)type, 4);
    if (len)
        crc = crc32(crc, data, len);
    unsigned char tail[4];
    put32(tail, (uint32_t)crc);
    return fwrite(hdr, 1, 8, f) == 8 && (!len || fwrite(data, 1, len, f) == len) &&
           fwrite(tail, 1, 4, f) == 4;
}

static unsigned char paeth(unsigned char a, unsigned char b, unsigned char c)
{
    const int p = (int)a + b - c;
    const int pa = abs(p - a), pb = abs(p - b), pc = abs(p - c);
    if (pa <= pb && pa <= pc)
        return a;
    return pb <= pc ? b : c;
}

/* filtered row (filter byte + n bytes) with the smallest sum of |signed byte| */
static void filter_row(const unsigned char* cur, const unsigned char* prev, size_t n,
                       unsigned char* cand, unsigned char* out)
{
    unsigned long best = ~0ul;
    for (int ft = 0; ft < 5; ++ft) {
        unsigned char* o = cand + (size_t)ft * (n + 1);
        o[0] = (unsigned char)ft;
        unsigned long sum = 0;
        for (size_t i = 0; i < n; ++i) {
            const unsigned char a = i >= 3 ? cur[i - 3] : 0;
            const unsigned char b = prev ? prev[i] : 0;
            const unsigned char c = (prev && i >= 3) ? prev[i - 3] : 0;
            unsigned char v = cur[i];
            switch (ft) {
            case 1: v = (unsigned char)(v - a); break;
            case 2: v = (unsigned char)(v - b); break;
            case 3: v = (unsigned char)(v - ((a + b) >> 1)); break;
            case 4: v = (unsigned char)(v - paeth(a, b, c)); break;
            default: break;
            }
            o[i + 1] = v;
            sum += v < 128 ? v : 256u - v;
        }
        if (sum < best) {
            best = sum;
            memcpy(out, o, n + 1);
        }
    }
}

int mdzg_png_write(const char* path, const uint32_t* rgb, int w, int h)
{
    if (!path || !rgb || w < 1 || h < 1 || (uint64_t)w * 3 + 1 > 0x7fffffffull)
        return MDZG_ERR_ARG;
    const size_t n = (size_t)w * 3;
    unsigned char* rows = (unsigned char*)malloc(2 * n);      /* current, previous */
    unsigned char* cand = (unsigned char*)malloc(5 * (n + 1));
    unsigned char* line = (unsigned char*)malloc(n + 1);
    const size_t zcap = 1 << 16;
    unsigned char* zbuf = (unsigned char*)malloc(zcap);
    FILE* f = fopen(path, "wb");
    int rc = (rows && cand && line && zbuf && f) ? MDZG_OK : MDZG_ERR_ARG;
    z_stream z;
    memset(&z, 0, sizeof z);
    int zinit = 0;
    if (rc == MDZG_OK) {
        static const unsigned char sig[8] = { 0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n' };
        unsigned char ihdr[13];
        put32(ihdr, (uint32_t)w);
        put32(ihdr + 4, (uint32_t)h);
        ihdr[8] = 8;                         /* bit depth            */
        ihdr[9] = 2;                         /* colour type RGB      */
        ihdr[10] = ihdr[11] = ihdr[12] = 0;  /* deflate, adaptive filtering, no interlace */
        if (fwrite(sig, 1, 8, f) != 8 || !chunk(f, "IHDR", ihdr, 13))
            rc = MDZG_ERR_ARG;
        if (deflateInit(&z, Z_DEFAULT_COMPRESSION) != Z_OK)
            rc = MDZG_ERR_NOMEM;
        else
            zinit = 1;
    }
    for (int y = 0; y <= h && rc == MDZG_OK; ++y) {
        const int last = y == h;
        if (!last) {
            unsigned char* cur = rows + (size_t)(y & 1) * n;
            const unsigned char* prev = y ? rows + (size_t)((y - 1) & 1) * n : NULL;
            const uint32_t* src = rgb + (size_t)y * (size_t)w;
            for (int x = 0; x < w; ++x) {
                cur[3 * x] = (unsigned char)src[x];
                cur[3 * x + 1] = (unsigned char)(src[x] >> 8);
                cur[3 * x + 2] = (unsigned char)(src[x] >> 16);
            }
            filter_row(cur, prev, n, cand, line);
            z.next_in = line;
            z.avail_in = (uInt)(n + 1);
        }
        int zr;
        do {
            z.next_out = zbuf;
            z.avail_out = (uInt)zcap;
            zr = deflate(&z, last ? Z_FINISH : Z_NO_FLUSH);
            if (zr == Z_STREAM_ERROR) {
                rc = MDZG_ERR_ARG;
                break;
            }
            const uint32_t got = (uint32_t)(zcap - z.avail_out);
            if (got && !chunk(f, "IDAT", zbuf, got)) {
                rc = MDZG_ERR_ARG;
                break;
            }
        } while (z.avail_out == 0 || (last && zr != Z_STREAM_END));
    }
    if (rc == MDZG_OK && !chunk(f, "IEND", NULL, 0))
        rc = MDZG_ERR_ARG;
    if (zinit)
        deflateEnd(&z);
    if (f && fclose(f) != 0 && rc == MDZG_OK)
        rc = MDZG_ERR_ARG;
    free(rows);
    free(cand);
    free(line);
    free(zbuf);
    return rc;
}
