/*
 * mdz_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).  See mdz_oracle.h.
 *
 * Built with -O2 -ffp-contract=off and no -march flags, so the `double`
 * restatement follows the C expression order with one SSE2 rounding per
 * operation, and the long double loops run on x87 with a 64-bit significand.
 *
 * Pinning (what checks each function; tests/test_oracle_*.py):
 *   ld_* / mpfr_* loops      : equal to the compiled reference loops
 *                              (oracle/_ref, built from src/frac_*.c) on the
 *                              same inputs, and the MPFR loops equal an
 *                              independent big-integer RN restatement
 *                              (tests/rn_ref.py) on small cases.
 *   d_* loops                : known counts from the closed forms of the
 *                              Mandelbrot set (points of the period-1/2 bulbs
 *                              never escape, |c|>2 escapes at iteration 1),
 *                              and agreement with the ld loop away from the
 *                              boundary.
 *   line workers (c chain)   : every c value equals the rounding chain redone
 *                              with mpmath at the same precision.
 *   rect_from_center         : mpmath restatement of the same MPFR chain.
 */
#define _GNU_SOURCE
#include "mdz_oracle.h"

#include <dlfcn.h>
#include <gmp.h>
#include <math.h>
#include <mpfr.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef int64_t depth_t;

/* ------------------------------------------------------------------ */
/* long double loops: src/frac_mandel.c:5-21, frac_burning_ship.c:7-24, */
/* frac_generalized_celtic.c:7-24, frac_variant.c:5-24                  */
/* ------------------------------------------------------------------ */
typedef depth_t (*ld_fn)(depth_t, long double, long double, long double,
                         long double, long double, long double);
typedef depth_t (*gmp_fn)(depth_t, mpf_t, mpf_t, mpf_t, mpf_t, mpf_t, mpf_t, mpf_t, mpf_t);
typedef depth_t (*mp_fn)(depth_t, mpfr_t, mpfr_t, mpfr_t, mpfr_t, mpfr_t,
                         mpfr_t, mpfr_t, mpfr_t);

static depth_t ld_mandel(depth_t depth, long double wim, long double wre,
                         long double c_im, long double c_re,
                         long double wim2, long double wre2)
{
    for (depth_t wz = 1; wz <= depth; ++wz) {
        wim = 2.0 * wre * wim + c_im;
        wre = wre2 - wim2 + c_re;
        wim2 = wim * wim;
        wre2 = wre * wre;
        if (wim2 + wre2 > 4.0F)
            return wz;
    }
    return 0;
}

static depth_t ld_ship(depth_t depth, long double wim, long double wre,
                       long double c_im, long double c_re,
                       long double wim2, long double wre2)
{
    for (depth_t wz = 1; wz <= depth; ++wz) {
        wim = 2.0 * fabsl(wre * wim) + c_im;
        wre = wre2 - wim2 + c_re;
        wim2 = wim * wim;
        wre2 = wre * wre;
        if (wim2 + wre2 > 4.0F)
            return wz;
    }
    return 0;
}

static depth_t ld_celtic(depth_t depth, long double wim, long double wre,
                         long double c_im, long double c_re,
                         long double wim2, long double wre2)
{
    for (depth_t wz = 1; wz <= depth; ++wz) {
        wim = 2.0 * wre * wim + c_im;
        wre = fabsl(wre2 - wim2) + c_re;
        wim2 = wim * wim;
        wre2 = wre * wre;
        if (wim2 + wre2 > 4.0F)
            return wz;
    }
    return 0;
}

static depth_t ld_variant(depth_t depth, long double wim, long double wre,
                          long double c_im, long double c_re,
                          long double wim2, long double wre2)
{
    for (depth_t wz = 1; wz <= depth; ++wz) {
        wim = 2.0 * wre * wim + c_im;
        if (wz % 2)
            wre = fabsl(wre2 - wim2) + c_re;
        else
            wre = wre2 - wim2 + c_re;
        wim2 = wim * wim;
        wre2 = wre * wre;
        if (wim2 + wre2 > 4.0F)
            return wz;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* double restatement of the same loops (the "CPU double loop")         */
/* ------------------------------------------------------------------ */
static depth_t d_loop(int fractal, depth_t depth, double wim, double wre,
                      double c_im, double c_re, double wim2, double wre2)
{
    for (depth_t wz = 1; wz <= depth; ++wz) {
        if (fractal == ORC_BURNING_SHIP)
            wim = 2.0 * fabs(wre * wim) + c_im;
        else
            wim = 2.0 * wre * wim + c_im;
        if (fractal == ORC_GENERALIZED_CELTIC ||
            (fractal == ORC_VARIANT && (wz % 2)))
            wre = fabs(wre2 - wim2) + c_re;
        else
            wre = wre2 - wim2 + c_re;
        wim2 = wim * wim;
        wre2 = wre * wre;
        if (wim2 + wre2 > 4.0F)
            return wz;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* MPFR loops: src/frac_mandel.c:25-52 and siblings.  Per iteration:    */
/* t=RN(wre*wim) [BS: |t|]; t=2t; wim=RN(t+c_im); t=RN(wre2-wim2)       */
/* [Celtic, odd-wz Variant: |t|]; wre=RN(t+c_re); wim2=RN(wim^2);       */
/* wre2=RN(wre^2); escape when RN(wim2+wre2) > bail.                    */
/* ------------------------------------------------------------------ */
static depth_t mp_loop(int fractal, depth_t depth, mpfr_t bail,
                       mpfr_t wim, mpfr_t wre, mpfr_t c_im, mpfr_t c_re,
                       mpfr_t wim2, mpfr_t wre2, mpfr_t t1)
{
    for (depth_t wz = 1; wz <= depth; wz++) {
        mpfr_mul(t1, wre, wim, MPFR_RNDN);
        if (fractal == ORC_BURNING_SHIP)
            mpfr_abs(t1, t1, MPFR_RNDN);
        mpfr_mul_si(t1, t1, 2, MPFR_RNDN);
        mpfr_add(wim, t1, c_im, MPFR_RNDN);
        mpfr_sub(t1, wre2, wim2, MPFR_RNDN);
        if (fractal == ORC_GENERALIZED_CELTIC ||
            (fractal == ORC_VARIANT && (wz % 2)))
            mpfr_abs(t1, t1, MPFR_RNDN);
        mpfr_add(wre, t1, c_re, MPFR_RNDN);
        mpfr_mul(wim2, wim, wim, MPFR_RNDN);
        mpfr_mul(wre2, wre, wre, MPFR_RNDN);
        mpfr_add(t1, wim2, wre2, MPFR_RNDN);
        if (mpfr_greater_p(t1, bail))
            return wz;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* GMP loops: src/frac_mandel.c:55-80, src/frac_burning_ship.c:58-83,     */
/* src/frac_generalized_celtic.c:58-83, src/frac_variant.c:58-87.  The     */
/* same operation order in GMP's truncating mpf arithmetic (mpf_mul,       */
/* mpf_mul_ui, mpf_add, mpf_sub, mpf_abs, mpf_cmp of GMP 6.2.1).          */
/* ------------------------------------------------------------------ */
static depth_t gmp_loop(int fractal, depth_t depth, mpf_t bail,
                        mpf_t wim, mpf_t wre, mpf_t c_im, mpf_t c_re,
                        mpf_t wim2, mpf_t wre2, mpf_t t1)
{
    for (depth_t wz = 1; wz <= depth; wz++) {
        mpf_mul(t1, wre, wim);
        if (fractal == ORC_BURNING_SHIP)
            mpf_abs(t1, t1);
        mpf_mul_ui(t1, t1, 2);
        mpf_add(wim, t1, c_im);
        mpf_sub(t1, wre2, wim2);
        if (fractal == ORC_GENERALIZED_CELTIC || (fractal == ORC_VARIANT && (wz % 2)))
            mpf_abs(t1, t1);
        mpf_add(wre, t1, c_re);
        mpf_mul(wim2, wim, wim);
        mpf_mul(wre2, wre, wre);
        mpf_add(t1, wim2, wre2);
        if (mpf_cmp(t1, bail) > 0)
            return wz;
    }
    return 0;
}

/* compiled reference loops, when loaded (oracle/_ref/libfrac_ref.so) */
static void*  ref_handle = NULL;
static ld_fn  ref_ld[4];
static mp_fn  ref_mp[4];
static gmp_fn ref_gmp[4];

int orc_version(void) { return 1; }

int orc_use_reference(const char* so_path)
{
    static const char* ld_names[4] = { "frac_mandel", "frac_burning_ship",
                                       "frac_generalized_celtic", "frac_variant" };
    static const char* mp_names[4] = { "frac_mandel_mpfr", "frac_burning_ship_mpfr",
                                       "frac_generalized_celtic_mpfr",
                                       "frac_variant_mpfr" };
    static const char* gmp_names[4] = { "frac_mandel_gmp", "frac_burning_ship_gmp",
                                        "frac_generalized_celtic_gmp", "frac_variant_gmp" };
    if (!so_path) {
        memset(ref_ld, 0, sizeof ref_ld);
        memset(ref_mp, 0, sizeof ref_mp);
        memset(ref_gmp, 0, sizeof ref_gmp);
        return 0;
    }
    /* one handle per build of the loops (the -O0 and -O2 libraries); a
       handle stays open until the process ends */
    static char ref_loaded[4096];
    if (!ref_handle || strcmp(ref_loaded, so_path) != 0) {
        void* h = dlopen(so_path, RTLD_NOW | RTLD_LOCAL);
        if (!h)
            return -1;
        ref_handle = h;
        snprintf(ref_loaded, sizeof ref_loaded, "%s", so_path);
    }
    for (int i = 0; i < 4; ++i) {
        ref_ld[i] = (ld_fn)dlsym(ref_handle, ld_names[i]);
        ref_mp[i] = (mp_fn)dlsym(ref_handle, mp_names[i]);
        ref_gmp[i] = (gmp_fn)dlsym(ref_handle, gmp_names[i]);
        if (!ref_ld[i] || !ref_mp[i] || !ref_gmp[i]) {
            memset(ref_ld, 0, sizeof ref_ld);
            memset(ref_mp, 0, sizeof ref_mp);
            memset(ref_gmp, 0, sizeof ref_gmp);
            return -1;
        }
    }
    return 0;
}

static depth_t run_ld(int fractal, depth_t depth, long double wim, long double wre,
                      long double c_im, long double c_re,
                      long double wim2, long double wre2)
{
    static const ld_fn own[4] = { ld_mandel, ld_ship, ld_celtic, ld_variant };
    ld_fn f = ref_ld[0] ? ref_ld[fractal] : own[fractal];
    return f(depth, wim, wre, c_im, c_re, wim2, wre2);
}

static depth_t run_mp(int fractal, depth_t depth, mpfr_t bail, mpfr_t wim,
                      mpfr_t wre, mpfr_t c_im, mpfr_t c_re, mpfr_t wim2,
                      mpfr_t wre2, mpfr_t t1)
{
    if (ref_mp[0])
        return ref_mp[fractal](depth, bail, wim, wre, c_im, c_re, wim2, wre2, t1);
    return mp_loop(fractal, depth, bail, wim, wre, c_im, c_re, wim2, wre2, t1);
}

static depth_t run_gmp(int fractal, depth_t depth, mpf_t bail, mpf_t wim, mpf_t wre,
                       mpf_t c_im, mpf_t c_re, mpf_t wim2, mpf_t wre2, mpf_t t1)
{
    if (ref_gmp[0])
        return ref_gmp[fractal](depth, bail, wim, wre, c_im, c_re, wim2, wre2, t1);
    return gmp_loop(fractal, depth, bail, wim, wre, c_im, c_re, wim2, wre2, t1);
}

/* ------------------------------------------------------------------ */
/* Frame state shared by the workers                                    */
/* ------------------------------------------------------------------ */
typedef struct job {
    const orc_frame* f;
    int32_t* out;
    long hp;                 /* max(p, 80): precision of img->xmin etc. */
    mpfr_t xmin, xmax, ymax, width, jre, jim;
    mpf_t gxmin, gymax, gwidth;   /* GMP path: img->gxmin/gymax/gwidth */
    pthread_mutex_t lock;    /* guards next_row (render_threads.c:366-369) */
    int next_row;
} job;

/* fractal.c:120-257 for the selected columns of one row */
static void line_mpfr(job* J, int line, int32_t* out)
{
    const orc_frame* f = J->f;
    mpfr_prec_t p = f->precision;
    mpfr_t x, y, x2, y2, c_re, c_im, wre, wim, wre2, wim2, bail, width, img_rw,
           img_xmin, t1;
    mpfr_inits2(p, x, y, x2, y2, c_re, c_im, wre, wim, wre2, wim2, bail, width,
                img_rw, img_xmin, t1, (mpfr_ptr)0);

    mpfr_set_si(bail, 4, MPFR_RNDN);
    mpfr_set_si(img_rw, f->real_width, MPFR_RNDN);
    mpfr_set(img_xmin, J->xmin, MPFR_RNDN);
    mpfr_set(width, J->width, MPFR_RNDN);

    mpfr_div(t1, width, img_rw, MPFR_RNDN);            /* fractal.c:167 */
    mpfr_mul_si(t1, t1, line, MPFR_RNDN);              /* :169 */
    mpfr_sub(y, J->ymax, t1, MPFR_RNDN);               /* :170 */
    mpfr_mul(y2, y, y, MPFR_RNDN);                     /* :171 */

    for (int k = 0; k < f->ncols; ++k) {
        int ix = f->col0 + k;
        mpfr_si_div(t1, ix, img_rw, MPFR_RNDN);        /* :183 */
        mpfr_mul(x, t1, width, MPFR_RNDN);             /* :185 */
        mpfr_add(x, x, img_xmin, MPFR_RNDN);           /* :186 */
        mpfr_mul(x2, x, x, MPFR_RNDN);                 /* :188 */
        mpfr_set(wre, x, MPFR_RNDN);
        mpfr_set(wim, y, MPFR_RNDN);
        mpfr_set(wre2, x2, MPFR_RNDN);
        mpfr_set(wim2, y2, MPFR_RNDN);
        if (f->family == ORC_FAMILY_JULIA) {
            mpfr_set(c_re, J->jre, MPFR_RNDN);
            mpfr_set(c_im, J->jim, MPFR_RNDN);
        } else {
            mpfr_set(c_re, x, MPFR_RNDN);
            mpfr_set(c_im, y, MPFR_RNDN);
        }
        out[k] = (int32_t)run_mp(f->fractal, f->depth, bail, wim, wre, c_im,
                                 c_re, wim2, wre2, t1);
    }
    mpfr_clears(x, y, x2, y2, c_re, c_im, wre, wim, wre2, wim2, bail, width,
                img_rw, img_xmin, t1, (mpfr_ptr)0);
}

/* coords.c:13-18 (mpfr_to_gmp) with src/my_mpfr_to_str.c (MPFR >= 2.4):
   the decimal string "%.Re" of the MPFR value, read by mpf_set_str */
static void to_gmp(mpfr_t from, mpf_t to)
{
    char d0[4096 + 1];                  /* MAX_DP, src/my_mpfr_to_str.h:10 */
    mpfr_snprintf(d0, 4096, "%.Re", from);
    mpf_set_str(to, d0, 10);
}

/* fractal.c:260-397 for the selected columns of one row */
static void line_gmp(job* J, int line, int32_t* out)
{
    const orc_frame* f = J->f;
    mp_bitcnt_t p = (mp_bitcnt_t)f->precision;
    mpf_t x, y, x2, y2, c_re, c_im, wre, wim, wre2, wim2, bail, width, img_rw, img_xmin, t1;
    mpf_t* all[] = { &x, &y, &x2, &y2, &c_re, &c_im, &wre, &wim, &wre2, &wim2, &bail,
                     &width, &img_rw, &img_xmin, &t1 };
    for (size_t i = 0; i < sizeof all / sizeof all[0]; ++i)
        mpf_init2(*all[i], p);

    mpf_set_si(bail, 4);
    mpf_set_si(img_rw, f->real_width);
    mpf_set(img_xmin, J->gxmin);
    mpf_set(width, J->gwidth);

    mpf_div(t1, width, img_rw);                        /* fractal.c:307 */
    mpf_mul_ui(t1, t1, (unsigned long)line);           /* :309 */
    mpf_sub(y, J->gymax, t1);                          /* :310 */
    mpf_mul(y2, y, y);                                 /* :311 */

    for (int k = 0; k < f->ncols; ++k) {
        int ix = f->col0 + k;
        mpf_ui_div(t1, (unsigned long)ix, img_rw);     /* :323 */
        mpf_mul(x, t1, width);                         /* :325 */
        mpf_add(x, x, img_xmin);                       /* :326 */
        mpf_mul(x2, x, x);                             /* :328 */
        mpf_set(wre, x);
        mpf_set(wim, y);
        mpf_set(wre2, x2);
        mpf_set(wim2, y2);
        if (f->family == ORC_FAMILY_JULIA) {
            to_gmp(J->jre, c_re);                      /* :341-342 */
            to_gmp(J->jim, c_im);
        } else {
            mpf_set(c_re, x);
            mpf_set(c_im, y);
        }
        out[k] = (int32_t)run_gmp(f->fractal, f->depth, bail, wim, wre, c_im, c_re,
                                  wim2, wre2, t1);
    }
    for (size_t i = 0; i < sizeof all / sizeof all[0]; ++i)
        mpf_clear(*all[i]);
}

/* fractal.c:29-117 for the selected columns of one row */
static void line_ld(job* J, int line, int32_t* out)
{
    const orc_frame* f = J->f;
    int W = f->real_width;
    long double xmin = mpfr_get_ld(J->xmin, MPFR_RNDN);
    long double xmax = mpfr_get_ld(J->xmax, MPFR_RNDN);
    long double ymax = mpfr_get_ld(J->ymax, MPFR_RNDN);
    long double width = xmax - xmin;
    long double jre = 0, jim = 0;
    if (f->family == ORC_FAMILY_JULIA) {
        jre = mpfr_get_ld(J->jre, MPFR_RNDN);
        jim = mpfr_get_ld(J->jim, MPFR_RNDN);
    }
    long double y = ymax - (width / (long double)W) * (long double)line;
    long double y2 = y * y;
    for (int k = 0; k < f->ncols; ++k) {
        int ix = f->col0 + k;
        long double x = (ix / (long double)W) * width + xmin;
        long double x2 = x * x;
        long double c_re = x, c_im = y;
        if (f->family == ORC_FAMILY_JULIA) { c_re = jre; c_im = jim; }
        out[k] = (int32_t)run_ld(f->fractal, f->depth, y, x, c_im, c_re, y2, x2);
    }
}

/* the same chain in `double` (the double path the FP64 kernel follows) */
static void line_d(job* J, int line, int32_t* out)
{
    const orc_frame* f = J->f;
    int W = f->real_width;
    double xmin = mpfr_get_d(J->xmin, MPFR_RNDN);
    double xmax = mpfr_get_d(J->xmax, MPFR_RNDN);
    double ymax = mpfr_get_d(J->ymax, MPFR_RNDN);
    double width = xmax - xmin;
    double jre = 0, jim = 0;
    if (f->family == ORC_FAMILY_JULIA) {
        jre = mpfr_get_d(J->jre, MPFR_RNDN);
        jim = mpfr_get_d(J->jim, MPFR_RNDN);
    }
    double y = ymax - (width / (double)W) * (double)line;
    double y2 = y * y;
    for (int k = 0; k < f->ncols; ++k) {
        int ix = f->col0 + k;
        double x = (ix / (double)W) * width + xmin;
        double x2 = x * x;
        double c_re = x, c_im = y;
        if (f->family == ORC_FAMILY_JULIA) { c_re = jre; c_im = jim; }
        out[k] = (int32_t)d_loop(f->fractal, f->depth, y, x, c_im, c_re, y2, x2);
    }
}

/* render_threads.c:342-393: claim the next row under a mutex, render it */
static void* worker(void* arg)
{
    job* J = (job*)arg;
    const orc_frame* f = J->f;
    for (;;) {
        pthread_mutex_lock(&J->lock);
        int k = J->next_row++;
        pthread_mutex_unlock(&J->lock);
        if (k >= f->nlines)
            break;
        int line = f->line0 + k * f->line_step;
        int32_t* out = J->out + (size_t)k * f->ncols;
        if (f->path == ORC_PATH_MPFR)
            line_mpfr(J, line, out);
        else if (f->path == ORC_PATH_GMP)
            line_gmp(J, line, out);
        else if (f->path == ORC_PATH_LD)
            line_ld(J, line, out);
        else
            line_d(J, line, out);
    }
    return NULL;
}

static void put_hex(char* dst, int cap, mpfr_t v)
{
    mpfr_snprintf(dst, (size_t)cap, "%Ra", v);
}

static int set_str(mpfr_t v, const char* s)
{
    return (s && mpfr_set_str(v, s, 0, MPFR_RNDN) == 0) ? 0 : -1;
}

int orc_render(const orc_frame* f, int nthreads, int32_t* out, double* elapsed_s)
{
    if (!f || !out || f->fractal < 0 || f->fractal > 3 || f->real_width < 1 ||
        f->real_height < 1 || f->nlines < 0 || f->ncols < 0 || f->col0 < 0 ||
        f->col0 + f->ncols > f->real_width || f->depth < 1 ||
        f->path < 0 || f->path > ORC_PATH_GMP ||
        ((f->path == ORC_PATH_MPFR || f->path == ORC_PATH_GMP) && f->precision < MPFR_PREC_MIN))
        return -1;
    if (nthreads < 1)
        nthreads = 1;

    job J;
    memset(&J, 0, sizeof J);
    J.f = f;
    J.out = out;
    J.hp = f->precision < 80 ? 80 : f->precision;
    /* the GMP path reads the coordinates as coords holds them, at the coords
       precision (src/coords.c:204-231, read by coords_get_rect_gmp :317-324),
       and the Julia constant as image_info holds it: c_re at max(p, 80), c_im
       at 80 bits (image_info.c:63-64 inits both at 80; :271-272 change the
       precision of c_re twice and never that of c_im) */
    const int gmp = f->path == ORC_PATH_GMP;
    long cp = J.hp;
    if (gmp) {
        mpf_t test;
        mpf_init2(test, (mp_bitcnt_t)cp);
        if ((long)mpf_get_prec(test) > cp)
            cp = (long)mpf_get_prec(test);
        mpf_clear(test);
    }
    mpfr_inits2(gmp ? cp : J.hp, J.xmin, J.xmax, J.ymax, J.width, (mpfr_ptr)0);
    mpfr_init2(J.jre, J.hp);
    mpfr_init2(J.jim, gmp ? 80 : J.hp);
    int bad = set_str(J.xmin, f->xmin) | set_str(J.xmax, f->xmax) |
              set_str(J.ymax, f->ymax) | set_str(J.width, f->width);
    if (f->family == ORC_FAMILY_JULIA)
        bad |= set_str(J.jre, f->julia_re) | set_str(J.jim, f->julia_im);
    if (bad) {
        mpfr_clears(J.xmin, J.xmax, J.ymax, J.width, J.jre, J.jim, (mpfr_ptr)0);
        return -1;
    }
    if (gmp) {
        /* image_info.c:274-277 (mpf at max(p, 80)), coords.c:317-324 */
        mpf_init2(J.gxmin, (mp_bitcnt_t)J.hp);
        mpf_init2(J.gymax, (mp_bitcnt_t)J.hp);
        mpf_init2(J.gwidth, (mp_bitcnt_t)J.hp);
        to_gmp(J.xmin, J.gxmin);
        to_gmp(J.ymax, J.gymax);
        to_gmp(J.width, J.gwidth);
    }
    pthread_mutex_init(&J.lock, NULL);

    pthread_t* th = (pthread_t*)calloc((size_t)nthreads, sizeof(pthread_t));
    struct timespec t0, t1;
    clock_gettime(CLOCK_MONOTONIC, &t0);
    int started = 0;
    for (int i = 0; i < nthreads; ++i)
        if (pthread_create(&th[i], NULL, worker, &J) == 0)
            ++started;
    if (started == 0)
        worker(&J);
    for (int i = 0; i < started; ++i)
        pthread_join(th[i], NULL);
    clock_gettime(CLOCK_MONOTONIC, &t1);
    if (elapsed_s)
        *elapsed_s = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);

    free(th);
    pthread_mutex_destroy(&J.lock);
    if (gmp) {
        mpf_clear(J.gxmin);
        mpf_clear(J.gymax);
        mpf_clear(J.gwidth);
    }
    mpfr_clears(J.xmin, J.xmax, J.ymax, J.width, J.jre, J.jim, (mpfr_ptr)0);
    mpfr_free_cache();
    return 0;
}

/* coords.c:204-231 (precision), coords.c:265-313 (center_to_rect, get_rect) */
static int rect_from_center(const char* cx_s, const char* cy_s, const char* size_s,
                            long precision, int user_w, int user_h, int at_coords,
                            char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap);

int orc_rect_from_center(const char* cx_s, const char* cy_s, const char* size_s,
                         long precision, int user_w, int user_h,
                         char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap)
{
    return rect_from_center(cx_s, cy_s, size_s, precision, user_w, user_h, 0,
                            xmin_o, xmax_o, ymax_o, width_o, cap);
}

/* the same rect as coords holds it, at the coords precision (GMP path input) */
int orc_rect_from_center_coords(const char* cx_s, const char* cy_s, const char* size_s,
                                long precision, int user_w, int user_h,
                                char* xmin_o, char* xmax_o, char* ymax_o, char* width_o,
                                int cap)
{
    return rect_from_center(cx_s, cy_s, size_s, precision, user_w, user_h, 1,
                            xmin_o, xmax_o, ymax_o, width_o, cap);
}

static int rect_from_center(const char* cx_s, const char* cy_s, const char* size_s,
                            long precision, int user_w, int user_h, int at_coords,
                            char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap)
{
    if (user_w < 1 || user_h < 1 || precision < 2)
        return -1;
    long cp = precision < 80 ? 80 : precision;
    mpf_t test;
    mpf_init2(test, (mp_bitcnt_t)cp);
    long gp = (long)mpf_get_prec(test);
    mpf_clear(test);
    if (gp > cp)
        cp = gp;
    long hp = precision < 80 ? 80 : precision;
    double aspect = (double)user_w / user_h;

    mpfr_t cx, cy, size, xmin, xmax, ymin, ymax, width, height, tmp, o;
    mpfr_inits2(precision, cx, cy, size, (mpfr_ptr)0);
    if (set_str(cx, cx_s) | set_str(cy, cy_s) | set_str(size, size_s)) {
        mpfr_clears(cx, cy, size, (mpfr_ptr)0);
        return -1;
    }
    mpfr_inits2(cp, xmin, xmax, ymin, ymax, width, height, tmp, (mpfr_ptr)0);
    mpfr_init2(o, at_coords ? cp : hp);
    if (aspect > 1.0) {
        mpfr_set(width, size, MPFR_RNDN);
        mpfr_div_d(height, width, aspect, MPFR_RNDN);
        mpfr_div_ui(tmp, width, 2, MPFR_RNDN);
        mpfr_sub(xmin, cx, tmp, MPFR_RNDN);
        mpfr_add(xmax, xmin, width, MPFR_RNDN);
        mpfr_div_d(tmp, tmp, aspect, MPFR_RNDN);
        mpfr_sub(ymin, cy, tmp, MPFR_RNDN);
        mpfr_add(ymax, ymin, height, MPFR_RNDN);
    } else {
        mpfr_set(height, size, MPFR_RNDN);
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
    mpfr_clears(cx, cy, size, xmin, xmax, ymin, ymax, width, height, tmp, o, (mpfr_ptr)0);
    return 0;
}

long long orc_formula_mpfr(int fractal, long long depth, long precision,
                           const char* wim_s, const char* wre_s,
                           const char* cim_s, const char* cre_s,
                           const char* wim2_s, const char* wre2_s,
                           char* out_wre, char* out_wim, int cap)
{
    if (fractal < 0 || fractal > 3 || precision < 2)
        return -1;
    mpfr_t bail, wim, wre, c_im, c_re, wim2, wre2, t1;
    mpfr_inits2(precision, bail, wim, wre, c_im, c_re, wim2, wre2, t1, (mpfr_ptr)0);
    long long r = -1;
    if (!(set_str(wim, wim_s) | set_str(wre, wre_s) | set_str(c_im, cim_s) |
          set_str(c_re, cre_s) | set_str(wim2, wim2_s) | set_str(wre2, wre2_s))) {
        mpfr_set_si(bail, 4, MPFR_RNDN);
        r = run_mp(fractal, depth, bail, wim, wre, c_im, c_re, wim2, wre2, t1);
        if (out_wre) put_hex(out_wre, cap, wre);
        if (out_wim) put_hex(out_wim, cap, wim);
    }
    mpfr_clears(bail, wim, wre, c_im, c_re, wim2, wre2, t1, (mpfr_ptr)0);
    return r;
}

int orc_rn_op(int op, long precision, const char* a_s, const char* b_s, char* out, int cap)
{
    if (precision < 2 || op < 0 || op > 2)
        return -1;
    mpfr_t a, b, r;
    /* operands parsed exactly (enough bits for any test operand) */
    mpfr_inits2(4096, a, b, (mpfr_ptr)0);
    mpfr_init2(r, precision);
    int rc = -1;
    if (!(set_str(a, a_s) | set_str(b, b_s))) {
        if (op == 0) mpfr_add(r, a, b, MPFR_RNDN);
        else if (op == 1) mpfr_sub(r, a, b, MPFR_RNDN);
        else mpfr_mul(r, a, b, MPFR_RNDN);
        put_hex(out, cap, r);
        rc = 0;
    }
    mpfr_clears(a, b, r, (mpfr_ptr)0);
    return rc;
}

long long orc_formula_ld(int fractal, long long depth,
                         const char* wim, const char* wre,
                         const char* c_im, const char* c_re,
                         const char* wim2, const char* wre2)
{
    if (fractal < 0 || fractal > 3 || !wim || !wre || !c_im || !c_re || !wim2 || !wre2)
        return -1;
    /* strtold reads decimal and hex-float strings exactly into the x87 format */
    return run_ld(fractal, depth, strtold(wim, NULL), strtold(wre, NULL),
                  strtold(c_im, NULL), strtold(c_re, NULL),
                  strtold(wim2, NULL), strtold(wre2, NULL));
}

long long orc_formula_d(int fractal, long long depth, double wim, double wre,
                        double c_im, double c_re, double wim2, double wre2)
{
    if (fractal < 0 || fractal > 3)
        return -1;
    return d_loop(fractal, depth, wim, wre, c_im, c_re, wim2, wre2);
}

int orc_pixel_c(const orc_frame* f, int line, int ix, char* xo, char* yo, int cap)
{
    if (!f || line < 0 || ix < 0 || ix >= f->real_width || f->real_width < 1)
        return -1;
    long hp = f->precision < 80 ? 80 : f->precision;
    mpfr_t xmin, xmax, ymax, width;
    mpfr_inits2(hp, xmin, xmax, ymax, width, (mpfr_ptr)0);
    int rc = -1;
    if (set_str(xmin, f->xmin) | set_str(xmax, f->xmax) | set_str(ymax, f->ymax) |
        set_str(width, f->width))
        goto out;
    const int W = f->real_width;
    if (f->path == ORC_PATH_MPFR) {
        mpfr_prec_t p = f->precision;
        mpfr_t t1, x, y, rw, xm, wd;
        mpfr_inits2(p, t1, x, y, rw, xm, wd, (mpfr_ptr)0);
        mpfr_set_si(rw, W, MPFR_RNDN);
        mpfr_set(xm, xmin, MPFR_RNDN);
        mpfr_set(wd, width, MPFR_RNDN);
        mpfr_div(t1, wd, rw, MPFR_RNDN);
        mpfr_mul_si(t1, t1, line, MPFR_RNDN);
        mpfr_sub(y, ymax, t1, MPFR_RNDN);
        mpfr_si_div(t1, ix, rw, MPFR_RNDN);
        mpfr_mul(x, t1, wd, MPFR_RNDN);
        mpfr_add(x, x, xm, MPFR_RNDN);
        put_hex(xo, cap, x);
        put_hex(yo, cap, y);
        mpfr_clears(t1, x, y, rw, xm, wd, (mpfr_ptr)0);
    } else if (f->path == ORC_PATH_LD) {
        long double x0 = mpfr_get_ld(xmin, MPFR_RNDN), x1 = mpfr_get_ld(xmax, MPFR_RNDN);
        long double y1 = mpfr_get_ld(ymax, MPFR_RNDN), w = x1 - x0;
        long double y = y1 - (w / (long double)W) * (long double)line;
        long double x = (ix / (long double)W) * w + x0;
        snprintf(xo, (size_t)cap, "%La", x);
        snprintf(yo, (size_t)cap, "%La", y);
    } else {
        double x0 = mpfr_get_d(xmin, MPFR_RNDN), x1 = mpfr_get_d(xmax, MPFR_RNDN);
        double y1 = mpfr_get_d(ymax, MPFR_RNDN), w = x1 - x0;
        double y = y1 - (w / (double)W) * (double)line;
        double x = (ix / (double)W) * w + x0;
        snprintf(xo, (size_t)cap, "%a", x);
        snprintf(yo, (size_t)cap, "%a", y);
    }
    rc = 0;
out:
    mpfr_clears(xmin, xmax, ymax, width, (mpfr_ptr)0);
    return rc;
}
