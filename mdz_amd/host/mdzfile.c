/*
 * mdzfile.c -- MDZ's fractal settings files (.mdz) for the command-line path:
 * the settings section as image_info_load_settings reads it
 * (src/image_info.c:456-721), with the line reader of src/mdzfileio.c and the
 * field parsers of src/setting.c, and the rect MDZ renders from it.
 *
 * Kept from the reference, because files written by MDZ depend on it:
 *   - a header "mdz fractal settings [maj.min.rev]" (no version = 0.0.8);
 *   - lines starting with '#' are skipped; tabs read as spaces and a line
 *     ends at its first control character (mdzfileio.c:15-34);
 *   - the fields come in a fixed order, each on its own line; before 0.1.0
 *     "fractal" names the family and "mpfr" the multi-precision flag, with
 *     rounding always on (image_info.c:481-541);
 *   - an option field whose line names a different field reads as option 0,
 *     and numbers parse in the C locale (setting.c:56-90, :94-170);
 *   - precision >= 80 (image_info.c:543), depth in [1, INT32_MAX];
 *   - coordinates either cx/cy/size or xmin/xmax/ymax (image_info.c:555-590),
 *     decimal strings read by MPFR at the file's precision;
 *   - julia-real/julia-imag for the Julia family; an optional palette-offset.
 */
#include <ctype.h>
#include <locale.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdzg_host.h"

#define LINELEN 4096                     /* FBUFLEN, src/mdzfileio.h:4 */
#define LOAD_WIDTH 480                   /* DEFAULT_WIDTH, src/image_info.h:22 */
#define STALE_SIZE 4.0                   /* image_info_create's size, src/image_info.c:33 */

typedef struct reader {
    FILE* f;
    char line[LINELEN + 1];
    char back[LINELEN + 1];              /* a line read ahead and pushed back */
    int has_back;
} reader;

/* next non-comment line, tabs as spaces, cut at the first control character
   (or byte >= 0x80) */
static int next_line(reader* r)
{
    do {
        if (r->has_back) {
            memcpy(r->line, r->back, sizeof r->line);
            r->has_back = 0;
        } else {
            if (!fgets(r->line, LINELEN, r->f))
                return 0;
            for (char* p = r->line; *p; ++p) {
                if (*p == '\t') {
                    *p = ' ';
                } else if ((signed char)*p < ' ') {   /* bytes >= 0x80 too, as char is signed */
                    *p = '\0';
                    break;
                }
            }
        }
    } while (r->line[0] == '#');
    return 1;
}

/* src/setting.c:56-90: the option index, -1 for an unknown option, and 0 when
   the line names another field */
static int get_index(const char* buf, const char* name, const char* const* options)
{
    size_t nlen = 0;
    if (name) {
        nlen = strlen(name);
        if (strlen(buf) < nlen + 2 || strncmp(name, buf, nlen))
            return 0;
    }
    const char* opt = buf + nlen;
    while (*opt == ' ')
        ++opt;
    for (int i = 0; options[i]; ++i)
        if (strcmp(opt, options[i]) == 0)
            return i;
    return -1;
}

static int name_ok(const char* buf, const char* name)
{
    const size_t nlen = strlen(name);
    return !(strlen(buf) < nlen + 1 || strncmp(name, buf, nlen));
}

/* sscanf in the C locale, as setting.c:94-170 sets it */
static int scan_c(const char* s, const char* fmt, void* out)
{
    char* old = setlocale(LC_NUMERIC, NULL);
    char saved[64] = "C";
    if (old)
        snprintf(saved, sizeof saved, "%s", old);
    setlocale(LC_NUMERIC, "C");
    const int r = sscanf(s, fmt, out);
    setlocale(LC_NUMERIC, saved);
    return r;
}

static int get_long(const char* buf, const char* name, long* v, long lo, long hi)
{
    if (!name_ok(buf, name) || scan_c(buf + strlen(name), "%ld", v) != 1)
        return 0;
    return (!lo && !hi) || (*v >= lo && *v <= hi);
}

static int get_double(const char* buf, const char* name, double* v, double lo, double hi)
{
    if (!name_ok(buf, name) || scan_c(buf + strlen(name), "%lf", v) != 1)
        return 0;
    return (!lo && !hi) || (*v >= lo && *v <= hi);
}

/* the value string of `name` if MPFR reads it as a base-10 number */
static char* get_number(const char* buf, const char* name)
{
    if (!name_ok(buf, name))
        return NULL;
    const char* s = buf + strlen(name);
    mpfr_t t;
    mpfr_init2(t, 64);
    const int bad = mpfr_set_str(t, s, 10, MPFR_RNDN);
    mpfr_clear(t);
    if (bad)
        return NULL;
    while (*s == ' ')
        ++s;
    return strdup(s);
}

/* mdzfile_test_for_name (src/mdzfileio.c:66-86): the line names `name` and
   has room for a value */
static int test_name(const char* buf, const char* name)
{
    const size_t nlen = strlen(name);
    return !(strlen(buf) < nlen + 2 || strncmp(name, buf, nlen));
}

/* palette_read (src/palette.c:187-217): up to 256 lines " r g b" read raw,
   stopping at the first that is not three integers; the colour count, 0 if
   none.  RGB(r, g, b) with two's complement shifts, as MDZ builds it. */
static int palette_read(FILE* f, uint32_t* colours)
{
    char buf[256];                                   /* BUF_SIZE, src/palette.c:8 */
    int i = 0;
    while (fgets(buf, sizeof buf, f)) {
        if (i >= 256)
            break;
        int rr, gg, bb;
        if (sscanf(buf, " %d %d %d", &rr, &gg, &bb) != 3)
            break;
        colours[i++] = (uint32_t)rr | ((uint32_t)gg << 8) | ((uint32_t)bb << 16);
    }
    return i;
}

int mdzg_palette_load(const char* path, mdzg_palette* pal)
{
    if (!path || !pal)
        return MDZG_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f)
        return MDZG_ERR_ARG;
    uint32_t c[256];
    const int n = palette_read(f, c);
    fclose(f);
    if (n == 0)
        return MDZG_ERR_PARSE;
    memcpy(pal->colours, c, sizeof(uint32_t) * (size_t)n);
    pal->n = n;
    return MDZG_OK;
}

int mdzg_scene_palette(const mdzg_scene* s, mdzg_palette* pal)
{
    if (!s || !pal)
        return MDZG_ERR_ARG;
    if (s->palette_n < 1)
        return MDZG_ERR_PARSE;
    memset(pal, 0, sizeof *pal);
    memcpy(pal->colours, s->palette, sizeof(uint32_t) * (size_t)s->palette_n);
    pal->n = s->palette_n;
    pal->offset = s->palette_offset;
    pal->colour_scale = s->colour_scale;
    pal->interpolate = s->colour_interpolate;
    return MDZG_OK;
}

static const char* const opt_no_yes[] = { "no", "yes", NULL };
static const char* const family_str[] = { "mandelbrot", "julia", NULL };       /* fractal.c:11-16 */
static const char* const fractal_str[] = { "mandelbrot", "burning ship", "generalized celtic",
                                           "mandel-celtic hybrid", NULL };     /* fractal.c:19-26 */
static const char* const coords_str[] = { "cx", "xmin", NULL };              /* image_info.c:24 */

void mdzg_scene_free(mdzg_scene* s)
{
    if (!s)
        return;
    free(s->c0);
    free(s->c1);
    free(s->c2);
    free(s->julia_re);
    free(s->julia_im);
    memset(s, 0, sizeof *s);
}

int mdzg_scene_load(const char* path, mdzg_scene* s)
{
    if (!path || !s)
        return MDZG_ERR_ARG;
    memset(s, 0, sizeof *s);
    reader r;
    memset(&r, 0, sizeof r);
    r.f = fopen(path, "r");
    if (!r.f)
        return MDZG_ERR_ARG;
    int rc = MDZG_ERR_PARSE;
    static const char header[] = "mdz fractal settings";
    if (!next_line(&r) || strncmp(r.line, header, sizeof header - 1))
        goto out;
    {
        const char* ver = r.line + sizeof header - 1;                /* mdzfileio.c:161-176 */
        while (*ver == ' ')
            ++ver;
        if (!*ver)
            ver = "0.0.8";
        int a, b, c;
        char* old = setlocale(LC_NUMERIC, NULL);
        char saved[64] = "C";
        if (old)
            snprintf(saved, sizeof saved, "%s", old);
        setlocale(LC_NUMERIC, "C");
        const int n = sscanf(ver, "%d.%d.%d", &a, &b, &c);
        setlocale(LC_NUMERIC, saved);
        if (n != 3)
            goto out;
        s->version_major = a;
        s->version_minor = b;
        s->version_revision = c;
    }
    if (!next_line(&r))                   /* the section line (image_info.c:753) */
        goto out;
    const int old_format = s->version_major == 0 && s->version_minor == 0;
    if (old_format) {
        if (!next_line(&r) || (s->family = get_index(r.line, "fractal", family_str)) == -1)
            goto out;
        s->fractal = MDZG_MANDELBROT;
    } else {
        if (!next_line(&r) || (s->family = get_index(r.line, "family", family_str)) == -1)
            goto out;
        if (!next_line(&r) || (s->fractal = get_index(r.line, "fractal", fractal_str)) == -1)
            goto out;
    }
    long depth;
    if (!next_line(&r) || !get_long(r.line, "depth", &depth, 1, 2147483647L))
        goto out;
    s->depth = depth;
    if (!next_line(&r) || !get_double(r.line, "aspect", &s->aspect, 1e-20f, 1e20f))
        goto out;
    if (!next_line(&r) || !get_double(r.line, "colour-scale", &s->colour_scale, 1e-20f, 1e20f))
        goto out;
    if (!next_line(&r) ||
        (s->colour_interpolate = get_index(r.line, "colour-interpolate", opt_no_yes)) == -1)
        goto out;
    if (old_format) {
        if (!next_line(&r) || (s->multi_precision = get_index(r.line, "mpfr", opt_no_yes)) == -1)
            goto out;
        s->rounding = 1;                  /* old versions always used MPFR */
    } else {
        if (!next_line(&r) ||
            (s->multi_precision = get_index(r.line, "multi-precision", opt_no_yes)) == -1)
            goto out;
        if (!next_line(&r) || (s->rounding = get_index(r.line, "multi-rounding", opt_no_yes)) == -1)
            goto out;
    }
    /* image_info_set_multi_prec (src/image_info.c:236-246) */
    s->path = s->multi_precision ? (s->rounding ? MDZG_PATH_MPFR : MDZG_PATH_GMP) : MDZG_PATH_LD;
    if (!next_line(&r) || !get_long(r.line, "precision", &s->precision, 80, 99999999))
        goto out;
    /* coordinates: the first word of the line names the form (image_info.c:555) */
    if (!next_line(&r))
        goto out;
    {
        char word[LINELEN + 1];
        const char* p = r.line;
        while (*p == ' ')
            ++p;
        size_t n = 0;
        while (p[n] && p[n] != ' ')
            ++n;
        memcpy(word, r.line, (size_t)(p - r.line) + n);
        word[(p - r.line) + n] = '\0';
        const int form = get_index(word, NULL, coords_str);
        if (form == -1)
            goto out;
        s->centre = form == 0;
    }
    if (!(s->c0 = get_number(r.line, s->centre ? "cx" : "xmin")))
        goto out;
    if (!next_line(&r) || !(s->c1 = get_number(r.line, s->centre ? "cy" : "xmax")))
        goto out;
    if (!next_line(&r) || !(s->c2 = get_number(r.line, s->centre ? "size" : "ymax")))
        goto out;
    if (s->family == MDZG_FAMILY_JULIA) {
        if (!next_line(&r) || !(s->julia_re = get_number(r.line, "julia-real")))
            goto out;
        if (!next_line(&r) || !(s->julia_im = get_number(r.line, "julia-imag")))
            goto out;
    }
    /* optional palette offset and random-palette settings (image_info.c:601-663),
       each tested on a line read ahead (mdzfile_test_for_name) */
    int have = next_line(&r);
    if (have && test_name(r.line, "palette-offset")) {
        long off;
        if (!get_long(r.line, "palette-offset", &off, 0, 255))
            goto out;
        s->palette_offset = (int)off;
        have = next_line(&r);
    }
    if (have && test_name(r.line, "r-strength")) {
        static const char* const rnd_d[] = { "r-strength", "r-bands", "g-strength", "g-bands",
                                             "b-strength", "b-bands" };
        static const char* const rnd_l[] = { "rnd-offset", "rnd-stripe", "rnd-spread" };
        for (int i = 0; i < 6; ++i) {
            double v;
            if ((i && !next_line(&r)) || !get_double(r.line, rnd_d[i], &v, 0.0, 1.0))
                goto out;
        }
        for (int i = 0; i < 3; ++i) {
            long v;
            if (!next_line(&r) || !get_long(r.line, rnd_l[i], &v, 0, 255))
                goto out;
        }
        have = next_line(&r);
    }
    /* the palette section (image_info_load_all, image_info.c:769-783, and
       image_info_load_palette, :724-740): "palette", then "data" and the
       colours, or "file <path>".  MDZ's command line refuses a file whose
       palette is missing or unreadable; here the settings still load, with
       palette_n = 0 */
    while (have && strcmp(r.line, "palette") != 0)
        have = next_line(&r);
    if (have && next_line(&r)) {
        if (strcmp(r.line, "data") == 0) {
            s->palette_n = palette_read(r.f, s->palette);
        } else if (test_name(r.line, "file")) {
            const char* fn = r.line + strlen("file");
            while (*fn == ' ')
                ++fn;
            FILE* pf = fopen(fn, "r");
            if (pf) {
                s->palette_n = palette_read(pf, s->palette);
                fclose(pf);
            }
        }
    }
    rc = MDZG_OK;
out:
    fclose(r.f);
    if (rc != MDZG_OK)
        mdzg_scene_free(s);
    return rc;
}

static void put_hex(char* dst, int cap, mpfr_srcptr v)
{
    mpfr_snprintf(dst, (size_t)cap, "%Ra", v);
}

/* the rect MDZ renders a scene at user_w x user_h output */
/* the image size MDZ's command line renders a scene at (main.c init_misc):
   a missing width is the loaded 480, a missing height follows the aspect
   image_info_set stored at load time, 480 / (int)(480 / file aspect) */
int mdzg_scene_dims(const mdzg_scene* s, int* w, int* h)
{
    if (!s || !w || !h || *w < 0 || *h < 0 || !(s->aspect > 0.0))
        return MDZG_ERR_ARG;
    const int h0 = (int)(LOAD_WIDTH / s->aspect);
    if (h0 < 1)
        return MDZG_ERR_ARG;
    const double aspect0 = (double)LOAD_WIDTH / h0;
    if (!*w && !*h) {
        *w = LOAD_WIDTH;
        *h = h0;
    } else if (!*w) {
        *w = (int)(*h * aspect0);
    } else if (!*h) {
        *h = (int)(*w / aspect0);
    }
    return (*w >= 1 && *h >= 1) ? MDZG_OK : MDZG_ERR_ARG;
}

int mdzg_scene_rect(const mdzg_scene* s, int user_w, int user_h, int flags,
                    char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap)
{
    if (!s || !s->c0 || user_w < 1 || user_h < 1 || !xmin_o || !xmax_o || !ymax_o ||
        !width_o || cap < 8 || (flags & ~MDZG_SCENE_FIT_RECT) || !(s->aspect > 0.0))
        return MDZG_ERR_ARG;
    const int gmp = s->path == MDZG_PATH_GMP;
    if (s->centre)                        /* coords_to + coords_size (image_info.c:706-709) */
        return gmp ? mdzg_rect_from_center_coords(s->c0, s->c1, s->c2, s->precision, user_w,
                                                  user_h, xmin_o, xmax_o, ymax_o, width_o, cap)
                   : mdzg_rect_from_center(s->c0, s->c1, s->c2, s->precision, user_w, user_h,
                                           xmin_o, xmax_o, ymax_o, width_o, cap);
    /* coords_set_rect + coords_rect_to_center (src/coords.c:327-333, :234-252)
       at the aspect image_info_set gave coords while loading (480 wide,
       image_info.c:666-669) -- or, with MDZG_SCENE_FIT_RECT, at the output
       aspect -- then coords_get_rect's coords_center_to_rect (:265-313) at the
       output aspect */
    const int h0 = (int)(LOAD_WIDTH / s->aspect);
    if (h0 < 1)
        return MDZG_ERR_ARG;
    const long p = s->precision;
    long cp = p < 80 ? 80 : p;
    mpf_t test;
    mpf_init2(test, (mp_bitcnt_t)cp);
    if ((long)mpf_get_prec(test) > cp)
        cp = (long)mpf_get_prec(test);
    mpf_clear(test);
    const long hp = p < 80 ? 80 : p;
    const double aspect = (double)user_w / user_h;
    const double aspect0 = (flags & MDZG_SCENE_FIT_RECT) ? aspect : (double)LOAD_WIDTH / h0;
    mpfr_t a, b, c, xmin, xmax, ymin, ymax, width, height, cx, cy, tmp, o;
    mpfr_inits2(p, a, b, c, (mpfr_ptr)0);
    if (mpfr_set_str(a, s->c0, 10, MPFR_RNDN) | mpfr_set_str(b, s->c1, 10, MPFR_RNDN) |
        mpfr_set_str(c, s->c2, 10, MPFR_RNDN)) {
        mpfr_clears(a, b, c, (mpfr_ptr)0);
        return MDZG_ERR_PARSE;
    }
    mpfr_inits2(cp, xmin, xmax, ymin, ymax, width, height, cx, cy, tmp, (mpfr_ptr)0);
    mpfr_init2(o, gmp ? cp : hp);
    mpfr_set(xmin, a, MPFR_RNDN);
    mpfr_set(xmax, b, MPFR_RNDN);
    mpfr_set(ymax, c, MPFR_RNDN);
    mpfr_sub(width, xmax, xmin, MPFR_RNDN);
    mpfr_div_d(height, width, aspect0, MPFR_RNDN);
    mpfr_sub(ymin, ymax, height, MPFR_RNDN);
    mpfr_add(cx, xmin, xmax, MPFR_RNDN);
    mpfr_div_ui(cx, cx, 2, MPFR_RNDN);
    mpfr_add(cy, ymin, ymax, MPFR_RNDN);
    mpfr_div_ui(cy, cy, 2, MPFR_RNDN);
    if (!(flags & MDZG_SCENE_FIT_RECT)) {
        /* init_misc's image_info_set -> coords_set (src/coords.c:255-262) puts
           coords' _size back into the width (wide) or height (tall); the rect
           never updated _size, which still holds coords_reset's 4.0 */
        mpfr_set_d(aspect > 1.0 ? width : height, STALE_SIZE, MPFR_RNDN);
    }
    if (aspect > 1.0) {
        mpfr_div_d(height, width, aspect, MPFR_RNDN);
        mpfr_div_ui(tmp, width, 2, MPFR_RNDN);
        mpfr_sub(xmin, cx, tmp, MPFR_RNDN);
        mpfr_add(xmax, xmin, width, MPFR_RNDN);
        mpfr_div_d(tmp, tmp, aspect, MPFR_RNDN);
        mpfr_sub(ymin, cy, tmp, MPFR_RNDN);
        mpfr_add(ymax, ymin, height, MPFR_RNDN);
    } else {
        mpfr_mul_d(width, height, aspect, MPFR_RNDN);
        mpfr_div_ui(tmp, height, 2, MPFR_RNDN);
        mpfr_sub(ymin, cy, tmp, MPFR_RNDN);
        mpfr_add(ymax, ymin, height, MPFR_RNDN);
        mpfr_mul_d(tmp, tmp, aspect, MPFR_RNDN);
        mpfr_sub(xmin, cx, tmp, MPFR_RNDN);
        mpfr_add(xmax, xmin, width, MPFR_RNDN);
    }
    mpfr_set(o, xmin, MPFR_RNDN);  put_hex(xmin_o, cap, o);
    mpfr_set(o, xmax, MPFR_RNDN);  put_hex(xmax_o, cap, o);
    mpfr_set(o, ymax, MPFR_RNDN);  put_hex(ymax_o, cap, o);
    mpfr_set(o, width, MPFR_RNDN); put_hex(width_o, cap, o);
    mpfr_clears(a, b, c, xmin, xmax, ymin, ymax, width, height, cx, cy, tmp, o, (mpfr_ptr)0);
    return MDZG_OK;
}
