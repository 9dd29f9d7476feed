/*
 * mdz_colour.c -- TEST INFRASTRUCTURE ONLY (the oracle): counts to colours
 * as MDZ's command line computes them, on the CPU.
 *
 *   palette_apply     src/palette.c:406-463  (aa == 1, whole rows)
 *   get_pixel_colour  src/palette.c:466-502  (one sample, the anti-alias path)
 *   do_anti_aliasing  src/render.c:108-157   (aa x aa box average)
 *   RGB/RED/GREEN/BLUE src/palette.h:16-19   (little-endian layout)
 *
 * Compiled by gcc for x86-64 like MDZ itself, so the double arithmetic and
 * the double-to-integer conversions are the ones MDZ's own build performs.
 * palette.c and render.c need GTK headers, which this image lacks, so they
 * are restated, not compiled.  palette_apply's interpolating path indexes
 * the palette with (int)floor(v) and (int)ceil(v), which for v >= 2^31 is a
 * negative index (undefined behaviour in MDZ): orc_colour refuses such a
 * frame instead of reading outside the palette.
 */
#include <math.h>
#include <stdint.h>

#include "mdz_oracle.h"

#define O_RGB(r, g, b) ((uint32_t)(r) | ((uint32_t)(g) << 8) | ((uint32_t)(b) << 16))
#define O_R(x) ((x) & 0xffu)
#define O_G(x) (((x) >> 8) & 0xffu)
#define O_B(x) (((x) >> 16) & 0xffu)

/* the interpolation both paths share: weights ceil(v) - v and 1 - that on the
   palette entries at floor(v) and ceil(v), each channel truncated to 8 bits */
static uint32_t mix(const uint32_t* pal, uint32_t m, double v, uint32_t lo, uint32_t hi,
                    double cval)
{
    const double w_lo = cval - v;
    const double w_hi = 1.0 - w_lo;
    const uint32_t a = pal[lo % m], b = pal[hi % m];
    const uint8_t r = (uint8_t)(w_lo * O_R(a) + w_hi * O_R(b));
    const uint8_t g = (uint8_t)(w_lo * O_G(a) + w_hi * O_G(b));
    const uint8_t bl = (uint8_t)(w_lo * O_B(a) + w_hi * O_B(b));
    return O_RGB(r, g, bl);
}

/* palette_apply for one count; -1 outside MDZ's defined domain */
static int64_t colour_row_pixel(const uint32_t* pal, int n, int offset, double sc, int ip,
                                int raw)
{
    if (raw == 0)
        return 0;
    const double v = raw * sc + offset;
    const uint32_t m = (uint32_t)(n - 1);
    if (!ip)
        return pal[(uint32_t)v % m];
    if (!(v < 2147483647.0))
        return -1;
    const int hi = (int)ceil(v);
    const int lo = (int)floor(v);
    return mix(pal, m, v, (uint32_t)lo, (uint32_t)hi, (double)hi);
}

/* get_pixel_colour(val) */
static uint32_t colour_sample(const uint32_t* pal, int n, int offset, int ip, double val)
{
    const uint32_t m = (uint32_t)(n - 1);
    if (val != 0.0)
        val += offset;
    if (ip) {
        const double c = ceil(val);
        return mix(pal, m, val, (uint32_t)floor(val), (uint32_t)c, c);
    }
    if (val == 0.0)
        return 0;
    return pal[(uint32_t)val % m];
}

int orc_colour(const uint32_t* counts, int user_w, int user_h, int aa, const uint32_t* pal,
               int n, int offset, double sc, int ip, uint32_t* rgb)
{
    if (!counts || !pal || !rgb || user_w < 1 || user_h < 1 || aa < 1 || n < 2 || n > 256 ||
        offset < 0 || !(sc > 0.0))
        return -1;
    const int *raw_data = (const int*)counts;     /* image_info.raw_data is int */
    const int real_w = user_w * aa;
    for (int y = 0; y < user_h; ++y) {
        for (int x = 0; x < user_w; ++x) {
            uint32_t* out = &rgb[(int64_t)y * user_w + x];
            if (aa == 1) {
                const int64_t c = colour_row_pixel(pal, n, offset, sc, ip,
                                                   raw_data[(int64_t)y * real_w + x]);
                if (c < 0)
                    return -2;
                *out = (uint32_t)c;
                continue;
            }
            uint32_t r = 0, g = 0, b = 0;
            for (int yi = y * aa; yi < y * aa + aa; ++yi)
                for (int xi = x * aa; xi < x * aa + aa; ++xi) {
                    const uint32_t c = colour_sample(pal, n, offset, ip,
                                                     raw_data[(int64_t)yi * real_w + xi] * sc);
                    r += O_R(c);
                    g += O_G(c);
                    b += O_B(c);
                }
            const int aaaa = aa * aa;
            *out = O_RGB(r / aaaa, g / aaaa, b / aaaa);
        }
    }
    return 0;
}
