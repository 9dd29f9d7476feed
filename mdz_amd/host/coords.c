/*
 * coords.c -- host-side coordinate setup for the device kernels.
 *
 * The reference computes z0 = c per pixel inside the line workers; x depends
 * only on the column and y only on the row, so the host computes each once per
 * column and per row (O(W+H) MPFR operations per tile) and the device squares
 * them itself with the same rounding.
 *
 *   MPFR path   src/fractal.c:159-188: width_p = RN_p(img->width),
 *               xmin_p = RN_p(img->xmin), W_p = RN_p(real_width),
 *               y = RN_p(ymax - RN_p(RN_p(width_p / W_p) * line)),
 *               x = RN_p(RN_p(RN_p(ix / W_p) * width_p) + xmin_p),
 *               Julia c = RN_p(img->u.julia.c) (src/fractal.c:200-203).
 *   long double src/fractal.c:50-72: xmin/xmax/ymax = mpfr_get_ld(RNDN),
 *               width = xmax - xmin, the chain in x87 long double.
 *   double      the same chain in IEEE double (the FP64 kernel evaluates it
 *               per pixel from xmin, ymax, width).
 * The rect itself (xmin/xmax/ymax/width from centre and size) follows
 * src/coords.c:204-231 and :265-313.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdzg_host.h"

const int mdzg_nwords[MDZG_NCLASS_N] = MDZG_NWORDS_INIT;
const int mdzg_gmp_limbs[MDZG_NGMP] = MDZG_GMP_LIMBS_INIT;

/* GMP's limb precision of mpf_init2(x, p): __GMPF_BITS_TO_PREC (gmp.h) */
static long gmp_prec_limbs(long p)
{
    return ((p > 53 ? p : 53) + 2 * GMP_NUMB_BITS - 1) / GMP_NUMB_BITS;
}

int mdzg_class_of(int path, long precision, int* nw, int* kclr)
{
    if (path == MDZG_PATH_DOUBLE) {
        *nw = 0;
        *kclr = 0;
        return 0;
    }
    if (path == MDZG_PATH_GMP) {
        if (precision < 2 || precision > MDZG_MAX_PRECISION)
            return -1;
        const int lr = (int)gmp_prec_limbs(precision) + 1;   /* limbs a value may hold */
        for (int j = 0; j < MDZG_NGMP; ++j)
            if (mdzg_gmp_limbs[j] >= lr) {
                *nw = lr;
                *kclr = 0;
                return MDZG_KID_GMP0 + j;
            }
        return -1;
    }
    /* p >= 2: MPFR builds older than 4.0 have no 1-bit precision, and the
       device arithmetic is checked from 2 bits up */
    long p = (path == MDZG_PATH_LD) ? 64 : precision;
    if (p < 2 || p > MDZG_MAX_PRECISION)
        return -1;
    for (int i = 0; i < MDZG_NCLASS_N; ++i) {
        if (32L * mdzg_nwords[i] >= p) {
            *nw = mdzg_nwords[i];
            *kclr = (int)(32L * mdzg_nwords[i] - p);
            return 1 + 2 * i + (*kclr == 0 ? 1 : 0);
        }
    }
    return -1;
}

static int set_str(mpfr_t v, const char* s)
{
    return (s && mpfr_set_str(v, s, 0, MPFR_RNDN) == 0) ? 0 : -1;
}

int mdzg_view_parse(const mdzg_view* v, mdzg_pview* pv)
{
    memset(pv, 0, sizeof *pv);
    if (!v || v->fractal < 0 || v->fractal > 3 || v->family < 0 || v->family > 1 ||
        v->path < 0 || v->path > MDZG_PATH_GMP || v->real_width < 1 || v->real_height < 1 ||
        v->depth < 1 || v->depth > INT32_MAX || v->precision < 1)
        return MDZG_ERR_ARG;
    int nw, kclr;
    int kid = mdzg_class_of(v->path, v->precision, &nw, &kclr);
    if (kid < 0)
        return MDZG_ERR_RANGE;
    pv->v = *v;
    pv->v.xmin = pv->v.xmax = pv->v.ymax = pv->v.width = NULL;
    pv->v.julia_re = pv->v.julia_im = NULL;
    pv->kid = kid;
    pv->nw = nw;
    pv->kclr = kclr;
    pv->hp = v->precision < 80 ? 80 : v->precision;
    if (v->path == MDZG_PATH_DOUBLE)
        pv->vs = 0;
    else if (v->path == MDZG_PATH_GMP)
        pv->vs = MDZG_GF_VS(mdzg_gmp_limbs[kid - MDZG_KID_GMP0]);
    else
        pv->vs = MDZG_VS(nw);
    /* the GMP path reads coords' own rect at the coords precision
       (src/coords.c:204-231, :317-324) and image_info's Julia constant, whose
       c_im keeps the 80 bits of its initialisation (src/image_info.c:63-64,
       :271-272 change the precision of c_re twice and never that of c_im) */
    long cp = pv->hp;
    if (v->path == MDZG_PATH_GMP) {
        mpf_t test;
        mpf_init2(test, (mp_bitcnt_t)cp);
        if ((long)mpf_get_prec(test) > cp)
            cp = (long)mpf_get_prec(test);
        mpf_clear(test);
    }
    mpfr_inits2(cp, pv->xmin, pv->xmax, pv->ymax, pv->width, (mpfr_ptr)0);
    mpfr_init2(pv->jre, pv->hp);
    mpfr_init2(pv->jim, v->path == MDZG_PATH_GMP ? 80 : pv->hp);
    pv->parsed = 1;
    int bad = set_str(pv->xmin, v->xmin) | set_str(pv->xmax, v->xmax) |
              set_str(pv->ymax, v->ymax) | set_str(pv->width, v->width);
    if (v->family == MDZG_FAMILY_JULIA)
        bad |= set_str(pv->jre, v->julia_re) | set_str(pv->jim, v->julia_im);
    else {
        mpfr_set_zero(pv->jre, 1);
        mpfr_set_zero(pv->jim, 1);
    }
    if (bad) {
        mdzg_view_clear(pv);
        return MDZG_ERR_PARSE;
    }
    return MDZG_OK;
}

void mdzg_view_clear(mdzg_pview* pv)
{
    if (pv->parsed)
        mpfr_clears(pv->xmin, pv->xmax, pv->ymax, pv->width, pv->jre, pv->jim, (mpfr_ptr)0);
    pv->parsed = 0;
}

/* GMP-path stored value: limbs from the most significant (low word first),
   then size, exponent and sign (mpf_gmp.h gf<store>) */
void mdzg_store_gf(uint32_t* w, int store, mpf_srcptr x)
{
    memset(w, 0, sizeof(uint32_t) * (size_t)MDZG_GF_VS(store));
    const int n = x->_mp_size < 0 ? -x->_mp_size : x->_mp_size;
    for (int i = 0; i < n && i < store; ++i) {
        const mp_limb_t l = x->_mp_d[n - 1 - i];
        w[2 * i] = (uint32_t)l;
        w[2 * i + 1] = (uint32_t)(l >> 32);
    }
    w[2 * store] = (uint32_t)n;
    w[2 * store + 1] = (uint32_t)(int32_t)(n ? x->_mp_exp : 0);
    w[2 * store + 2] = x->_mp_size < 0 ? 1u : 0u;
}

/* src/coords.c:13-18 with src/my_mpfr_to_str.c: the decimal string "%.Re" of
   the MPFR value (its own precision decides the digits), read by mpf_set_str */
static void mpfr_to_mpf(mpfr_srcptr from, mpf_ptr to)
{
    char d0[4096 + 1];                     /* MAX_DP, src/my_mpfr_to_str.h:10 */
    mpfr_snprintf(d0, 4096, "%.Re", from);
    mpf_set_str(to, d0, 10);
}

/* ---- stored-value conversion ------------------------------------------ */
void mdzg_store_mpfr(uint32_t* w, int nw, mpfr_srcptr x)
{
    memset(w, 0, sizeof(uint32_t) * (size_t)MDZG_VS(nw));
    if (mpfr_zero_p(x)) {
        w[nw] = (uint32_t)(-(1 << 30));
        w[nw + 1] = 0;
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpfr_exp_t ex = mpfr_get_z_2exp(z, x);   /* x = z * 2^ex, |z| has prec bits */
    long prec = (long)mpfr_get_prec(x);
    mpz_abs(z, z);
    mpz_mul_2exp(z, z, (mp_bitcnt_t)(32L * nw - prec));
    size_t cnt = 0;
    mpz_export(w, &cnt, -1, 4, 0, 0, z);
    w[nw] = (uint32_t)(int32_t)(ex + prec);
    w[nw + 1] = mpfr_signbit(x) ? 1u : 0u;
    mpz_clear(z);
}

void mdzg_store_ld(uint32_t* w, int nw, long double x)
{
    memset(w, 0, sizeof(uint32_t) * (size_t)MDZG_VS(nw));
    if (x == 0.0L) {
        w[nw] = (uint32_t)(-(1 << 30));
        return;
    }
    int e = 0;
    long double m = frexpl(fabsl(x), &e);            /* m in [1/2, 1) */
    uint64_t M = (uint64_t)ldexpl(m, 64);            /* exact: 64-bit significand */
    w[nw - 2] = (uint32_t)M;
    w[nw - 1] = (uint32_t)(M >> 32);
    w[nw] = (uint32_t)(int32_t)e;
    w[nw + 1] = signbit(x) ? 1u : 0u;
}

/* ---- per-column / per-row values ---------------------------------------- */
int mdzg_coords_fill(const mdzg_pview* pv, int line0, int nlines, int col0, int ncols,
                     uint32_t* xs, uint32_t* ys, uint32_t* jc)
{
    const int nw = pv->nw;
    const int vs = pv->vs;
    const int W = pv->v.real_width;
    if (pv->v.path == MDZG_PATH_MPFR) {
        const mpfr_prec_t p = pv->v.precision;
        mpfr_t t1, x, y, img_rw, img_xmin, width;
        mpfr_inits2(p, t1, x, y, img_rw, img_xmin, width, (mpfr_ptr)0);
        mpfr_set_si(img_rw, W, MPFR_RNDN);
        mpfr_set(img_xmin, pv->xmin, MPFR_RNDN);
        mpfr_set(width, pv->width, MPFR_RNDN);
        mpfr_t step;
        mpfr_init2(step, p);
        mpfr_div(step, width, img_rw, MPFR_RNDN);                 /* fractal.c:167 */
        for (int k = 0; k < nlines; ++k) {
            mpfr_mul_si(t1, step, line0 + k, MPFR_RNDN);          /* :169 */
            mpfr_sub(y, pv->ymax, t1, MPFR_RNDN);                 /* :170 */
            mdzg_store_mpfr(ys + (size_t)k * vs, nw, y);
        }
        for (int k = 0; k < ncols; ++k) {
            mpfr_si_div(t1, col0 + k, img_rw, MPFR_RNDN);         /* :183 */
            mpfr_mul(x, t1, width, MPFR_RNDN);                    /* :185 */
            mpfr_add(x, x, img_xmin, MPFR_RNDN);                  /* :186 */
            mdzg_store_mpfr(xs + (size_t)k * vs, nw, x);
        }
        if (jc && pv->v.family == MDZG_FAMILY_JULIA) {
            mpfr_set(t1, pv->jre, MPFR_RNDN);                     /* :201-202 */
            mdzg_store_mpfr(jc, nw, t1);
            mpfr_set(t1, pv->jim, MPFR_RNDN);
            mdzg_store_mpfr(jc + vs, nw, t1);
        }
        mpfr_clears(t1, x, y, img_rw, img_xmin, width, step, (mpfr_ptr)0);
        return MDZG_OK;
    }
    if (pv->v.path == MDZG_PATH_GMP) {
        /* src/fractal.c:298-328 and image_info.c:274-277: gxmin/gymax/gwidth
           at max(p, 80), the chain at p, all in GMP's mpf arithmetic */
        const int store = mdzg_gmp_limbs[pv->kid - MDZG_KID_GMP0];
        const mp_bitcnt_t p = (mp_bitcnt_t)pv->v.precision;
        mpf_t gxmin, gymax, gwidth, t1, x, y, img_rw, img_xmin, width;
        mpf_init2(gxmin, (mp_bitcnt_t)pv->hp);
        mpf_init2(gymax, (mp_bitcnt_t)pv->hp);
        mpf_init2(gwidth, (mp_bitcnt_t)pv->hp);
        mpfr_to_mpf(pv->xmin, gxmin);
        mpfr_to_mpf(pv->ymax, gymax);
        mpfr_to_mpf(pv->width, gwidth);
        mpf_init2(t1, p);
        mpf_init2(x, p);
        mpf_init2(y, p);
        mpf_init2(img_rw, p);
        mpf_init2(img_xmin, p);
        mpf_init2(width, p);
        mpf_set_si(img_rw, W);
        mpf_set(img_xmin, gxmin);
        mpf_set(width, gwidth);
        for (int k = 0; k < nlines; ++k) {
            mpf_div(t1, width, img_rw);                           /* fractal.c:307 */
            mpf_mul_ui(t1, t1, (unsigned long)(line0 + k));       /* :309 */
            mpf_sub(y, gymax, t1);                                /* :310 */
            mdzg_store_gf(ys + (size_t)k * vs, store, y);
        }
        for (int k = 0; k < ncols; ++k) {
            mpf_ui_div(t1, (unsigned long)(col0 + k), img_rw);    /* :323 */
            mpf_mul(x, t1, width);                                /* :325 */
            mpf_add(x, x, img_xmin);                              /* :326 */
            mdzg_store_gf(xs + (size_t)k * vs, store, x);
        }
        if (jc && pv->v.family == MDZG_FAMILY_JULIA) {
            mpfr_to_mpf(pv->jre, t1);                             /* :341-342 */
            mdzg_store_gf(jc, store, t1);
            mpfr_to_mpf(pv->jim, t1);
            mdzg_store_gf(jc + vs, store, t1);
        }
        mpf_clear(gxmin);
        mpf_clear(gymax);
        mpf_clear(gwidth);
        mpf_clear(t1);
        mpf_clear(x);
        mpf_clear(y);
        mpf_clear(img_rw);
        mpf_clear(img_xmin);
        mpf_clear(width);
        return MDZG_OK;
    }
    if (pv->v.path == MDZG_PATH_LD) {
        long double xmin = mpfr_get_ld(pv->xmin, MPFR_RNDN);      /* fractal.c:50-53 */
        long double xmax = mpfr_get_ld(pv->xmax, MPFR_RNDN);
        long double ymax = mpfr_get_ld(pv->ymax, MPFR_RNDN);
        long double width = xmax - xmin;
        for (int k = 0; k < nlines; ++k) {
            long double y = ymax - (width / (long double)W) * (long double)(line0 + k);
            mdzg_store_ld(ys + (size_t)k * vs, nw, y);            /* :61-62 */
        }
        for (int k = 0; k < ncols; ++k) {
            long double x = ((col0 + k) / (long double)W) * width + xmin;   /* :72 */
            mdzg_store_ld(xs + (size_t)k * vs, nw, x);
        }
        if (jc && pv->v.family == MDZG_FAMILY_JULIA) {
            mdzg_store_ld(jc, nw, mpfr_get_ld(pv->jre, MPFR_RNDN)); /* :57-58 */
            mdzg_store_ld(jc + vs, nw, mpfr_get_ld(pv->jim, MPFR_RNDN));
        }
        return MDZG_OK;
    }
    return MDZG_ERR_ARG;
}

void mdzg_coords_fp64(const mdzg_pview* pv, double* xmin, double* ymax, double* width,
                      double* jre, double* jim)
{
    double x0 = mpfr_get_d(pv->xmin, MPFR_RNDN);
    double x1 = mpfr_get_d(pv->xmax, MPFR_RNDN);
    *xmin = x0;
    *ymax = mpfr_get_d(pv->ymax, MPFR_RNDN);
    *width = x1 - x0;
    *jre = mpfr_get_d(pv->jre, MPFR_RNDN);
    *jim = mpfr_get_d(pv->jim, MPFR_RNDN);
}

/* ---- rect from centre and size (command-line path) ----------------------- */
static void put_hex(char* dst, int cap, mpfr_srcptr v)
{
    mpfr_snprintf(dst, (size_t)cap, "%Ra", v);
}

static int rect_from_center(const char* cx_s, const char* cy_s, const char* size_s,
                            long precision, int user_w, int user_h, int at_coords,
                            char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap);

int mdzg_rect_from_center(const char* cx_s, const char* cy_s, const char* size_s,
                          long precision, int user_w, int user_h,
                          char* xmin_o, char* xmax_o, char* ymax_o, char* width_o, int cap)
{
    return rect_from_center(cx_s, cy_s, size_s, precision, user_w, user_h, 0,
                            xmin_o, xmax_o, ymax_o, width_o, cap);
}

int mdzg_rect_from_center_coords(const char* cx_s, const char* cy_s, const char* size_s,
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
    if (!cx_s || !cy_s || !size_s || !xmin_o || !xmax_o || !ymax_o || !width_o ||
        cap < 8 || user_w < 1 || user_h < 1 || precision < 2)
        return MDZG_ERR_ARG;
    /* coords precision: max(p, 80), raised to GMP's limb-granular mpf
       precision (src/coords.c:204-231) */
    long cp = precision < 80 ? 80 : precision;
    mpf_t test;
    mpf_init2(test, (mp_bitcnt_t)cp);
    long gp = (long)mpf_get_prec(test);
    mpf_clear(test);
    if (gp > cp)
        cp = gp;
    const long hp = precision < 80 ? 80 : precision;
    const double aspect = (double)user_w / user_h;

    mpfr_t cx, cy, size;
    mpfr_inits2(precision, cx, cy, size, (mpfr_ptr)0);      /* parsed at p */
    if (set_str(cx, cx_s) | set_str(cy, cy_s) | set_str(size, size_s)) {
        mpfr_clears(cx, cy, size, (mpfr_ptr)0);
        return MDZG_ERR_PARSE;
    }
    mpfr_t xmin, xmax, ymin, ymax, width, height, tmp, o;
    mpfr_inits2(cp, xmin, xmax, ymin, ymax, width, height, tmp, (mpfr_ptr)0);
    mpfr_init2(o, at_coords ? cp : hp);
    if (aspect > 1.0) {                                     /* coords.c:272-285 */
        mpfr_set(width, size, MPFR_RNDN);
        mpfr_div_d(height, width, aspect, MPFR_RNDN);
        mpfr_div_ui(tmp, width, 2, MPFR_RNDN);
        mpfr_sub(xmin, cx, tmp, MPFR_RNDN);
        mpfr_add(xmax, xmin, width, MPFR_RNDN);
        mpfr_div_d(tmp, tmp, aspect, MPFR_RNDN);
        mpfr_sub(ymin, cy, tmp, MPFR_RNDN);
        mpfr_add(ymax, ymin, height, MPFR_RNDN);
    } else {                                                /* coords.c:286-299 */
        mpfr_set(height, size, MPFR_RNDN);
        mpfr_mul_d(width, height, aspect, MPFR_RNDN);
        mpfr_div_ui(tmp, height, 2, MPFR_RNDN);
        mpfr_sub(ymin, cy, tmp, MPFR_RNDN);
        mpfr_add(ymax, ymin, height, MPFR_RNDN);
        mpfr_mul_d(tmp, tmp, aspect, MPFR_RNDN);
        mpfr_sub(xmin, cx, tmp, MPFR_RNDN);
        mpfr_add(xmax, xmin, width, MPFR_RNDN);
    }
    /* coords_get_rect: copied into image_info at max(p, 80) (coords.c:305-313);
       or, for the GMP path, as coords holds them (coords.c:317-324) */
    mpfr_set(o, xmin, MPFR_RNDN);  put_hex(xmin_o, cap, o);
    mpfr_set(o, xmax, MPFR_RNDN);  put_hex(xmax_o, cap, o);
    mpfr_set(o, ymax, MPFR_RNDN);  put_hex(ymax_o, cap, o);
    mpfr_set(o, width, MPFR_RNDN); put_hex(width_o, cap, o);
    mpfr_clears(cx, cy, size, xmin, xmax, ymin, ymax, width, height, tmp, o, (mpfr_ptr)0);
    return MDZG_OK;
}
