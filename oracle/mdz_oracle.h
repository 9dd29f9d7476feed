/*
 * mdz_oracle.h -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * CPU restatement of MDZ's per-pixel escape-time path, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product (mdz_amd/, include/) may link or call it.
 *
 * What it restates (reference file:line):
 *   formula loops, long double : src/frac_mandel.c:5-21, src/frac_burning_ship.c:7-24,
 *                                src/frac_generalized_celtic.c:7-24, src/frac_variant.c:5-24
 *   formula loops, MPFR        : src/frac_mandel.c:25-52, src/frac_burning_ship.c:27-55,
 *                                src/frac_generalized_celtic.c:27-55, src/frac_variant.c:26-55
 *   formula loops, double      : the long double loops with `double` (the "CPU double loop")
 *   formula loops, GMP (mpf)   : src/frac_mandel.c:55-80, src/frac_burning_ship.c:58-83,
 *                                src/frac_generalized_celtic.c:58-83, src/frac_variant.c:58-87
 *   line workers               : src/fractal.c:29-117 (long double), src/fractal.c:120-257 (MPFR),
 *                                src/fractal.c:260-397 (GMP), coords.c:13-18 (MPFR to mpf)
 *   rect from centre/size      : src/coords.c:204-231, :265-313, src/image_info.c:252-280
 *   line pool                  : src/render_threads.c:342-393 (dynamic line claim under a mutex)
 *   colouring                  : src/palette.c:406-502, src/render.c:108-157 (mdz_colour.c)
 *
 * The third-party arithmetic is MPFR (headers 4.0.2 under /opt/conda, GMP 6.2.1);
 * MPFR's RNDN results are fixed by its correct-rounding guarantee, so the
 * loops' bits do not depend on the MPFR version.
 */
#ifndef MDZ_ORACLE_H
#define MDZ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_MANDELBROT = 0, ORC_BURNING_SHIP = 1, ORC_GENERALIZED_CELTIC = 2, ORC_VARIANT = 3 };
enum { ORC_FAMILY_MANDEL = 0, ORC_FAMILY_JULIA = 1 };
enum { ORC_PATH_MPFR = 0, ORC_PATH_LD = 1, ORC_PATH_DOUBLE = 2, ORC_PATH_GMP = 3 };

typedef struct orc_frame {
    int fractal;            /* ORC_MANDELBROT .. ORC_VARIANT                      */
    int family;             /* ORC_FAMILY_MANDEL / ORC_FAMILY_JULIA               */
    int path;               /* ORC_PATH_*                                         */
    int real_width;         /* image_info.real_width  (user width  x aa)          */
    int real_height;        /* image_info.real_height (user height x aa)          */
    long precision;         /* image_info.precision (p of the MPFR loop)          */
    long long depth;        /* image_info.depth (max_iter)                        */
    /* image_info.xmin/xmax/ymax/width as MPFR-parsable strings (decimal or
       hex-float); they are rounded to max(p, 80) bits as image_info holds them.
       GMP path: the coords' own xmin/xmax/ymax/width, rounded to the coords
       precision (src/coords.c:204-231), which coords_get_rect_gmp converts. */
    const char* xmin;
    const char* xmax;
    const char* ymax;
    const char* width;
    const char* julia_re;   /* image_info.u.julia.c_re (Julia family only)        */
    const char* julia_im;
    /* which pixels: rows line0 + k*line_step (k < nlines), columns [col0, col0+ncols) */
    int line0, nlines, line_step;
    int col0, ncols;
} orc_frame;

int  orc_version(void);

/* Route the long double and MPFR formula loops through the compiled reference
   loops (oracle/_ref/libfrac_ref.so) instead of the restated ones.  NULL path
   switches back.  Returns 0 on success, -1 when the library cannot be loaded. */
int  orc_use_reference(const char* so_path);

/* Render the frame's selected pixels with `nthreads` workers claiming rows
   dynamically.  out: nlines*ncols int32 counts (0 = never escaped).
   Returns 0, or -1 on bad arguments.  If `elapsed_s` is non-NULL it receives
   the wall time from worker spawn to join. */
int  orc_render(const orc_frame* f, int nthreads, int32_t* out, double* elapsed_s);

/* Rect of a centre/size view as the command-line path computes it
   (coords_set_precision + coords_center_to_rect + coords_get_rect).
   Outputs hex-float strings of img->xmin/xmax/ymax/width, each at
   max(precision, 80) bits, into buffers of `cap` bytes. */
int  orc_rect_from_center(const char* cx, const char* cy, const char* size,
                          long precision, int user_w, int user_h,
                          char* xmin, char* xmax, char* ymax, char* width, int cap);
/* the same rect at the coords precision (what coords_get_rect_gmp reads) */
int  orc_rect_from_center_coords(const char* cx, const char* cy, const char* size,
                                 long precision, int user_w, int user_h,
                                 char* xmin, char* xmax, char* ymax, char* width, int cap);

/* One formula loop on explicit start values (strings, rounded to p).
   Returns the count; writes the final wre/wim as hex strings if non-NULL. */
long long orc_formula_mpfr(int fractal, long long depth, long precision,
                           const char* wim, const char* wre,
                           const char* c_im, const char* c_re,
                           const char* wim2, const char* wre2,
                           char* out_wre, char* out_wim, int cap);

/* The long double loop on explicit start values (x87, 64-bit significand),
   given as strings that strtold converts exactly (hex-float or decimal). */
long long orc_formula_ld(int fractal, long long depth,
                         const char* wim, const char* wre,
                         const char* c_im, const char* c_re,
                         const char* wim2, const char* wre2);

/* The double loop on explicit start values. */
long long orc_formula_d(int fractal, long long depth, double wim, double wre,
                        double c_im, double c_re, double wim2, double wre2);

/* c of pixel (line, ix) on the frame's path, as the line worker computes it:
   MPFR path hex strings at p bits; long double / double paths as "%La" / "%a". */
int orc_pixel_c(const orc_frame* f, int line, int ix, char* x, char* y, int cap);

/* Single RN operations at precision p on string operands (hex-float output),
   for per-operation checks: op 0 add, 1 sub, 2 mul. */
int  orc_rn_op(int op, long precision, const char* a, const char* b, char* out, int cap);

/* Colours of a frame of counts ((user_w*aa) x (user_h*aa) samples) as MDZ's
   command line computes them: palette_apply for aa == 1, do_anti_aliasing with
   get_pixel_colour for aa > 1.  0, -1 on bad arguments, -2 when palette_apply's
   interpolating path would index outside the palette (v >= 2^31). */
int orc_colour(const uint32_t* counts, int user_w, int user_h, int aa, const uint32_t* pal,
               int n, int offset, double sc, int ip, uint32_t* rgb);

#ifdef __cplusplus
}
#endif
#endif
