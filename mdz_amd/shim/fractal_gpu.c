/*
 * fractal_gpu.c -- the shim MDZ adds to route its line callback to the
 * MI355X engine (the file MDZ would compile as src/fractal_gpu.c).
 *
 * It has the shape of MDZ's line workers, int (*)(image_info*, int line)
 * (src/render_threads.h:70-71, src/fractal.h:42-44), and is installed where
 * MDZ picks its line callback (src/image_info.c:242-248):
 *
 *     if (use_multi_prec)
 *         rth_set_next_line_cb(rth, use_rounding ? fractal_gpu_mpfr_calculate_line
 *                                                : fractal_gpu_gmp_calculate_line);
 *     else
 *         rth_set_next_line_cb(rth, fractal_gpu_ld_calculate_line);
 *
 * MDZ's worker threads claim rows and call the callback once per row,
 * concurrently (src/render_threads.c:342-393).  The shim converts the
 * image_info MPFR fields to exact hex strings once per render and opens one
 * frame per render.  A render hands out every row exactly once and ends
 * early only on a 0 return, so the shim counts the rows it has handed out:
 * the first call after all real_height rows of the open frame were handed
 * out, or after a stop, belongs to a new render.  Row order does not
 * matter.  MDZ starts the next render only after it joined the previous
 * render's threads (src/render_threads.c:324-325), so no thread is still in
 * the old frame when it is closed.
 *
 * RT_STOP (rth_render_should_stop, src/render_threads.c:409-418) becomes
 * mdzg_frame_stop, which raises the device cancel flag of every band in
 * flight; rows never rendered keep the UINT_MAX sentinel MDZ filled raw_data
 * with (src/image_info.c:214-215).
 *
 * Build with MDZ's headers (fractal.h, render_threads.h) and include/mdzg.h,
 * link -lmdzg.
 */
#include <pthread.h>
#include <stdio.h>

#include <mdzg.h>

#include "fractal.h"
#include "render_threads.h"

static mdzg_ctx* gpu_ctx;
static mdzg_frame* gpu_frame;
static int gpu_rows_left;         /* rows of gpu_frame not yet handed out */
static pthread_mutex_t gpu_lock = PTHREAD_MUTEX_INITIALIZER;

static int open_frame(image_info* img, int path)
{
    char *xmin = NULL, *xmax = NULL, *ymax = NULL, *width = NULL, *jre = NULL, *jim = NULL;
    /* the GMP path reads coords' own rect (coords_get_rect_gmp, src/render.c:36-37) */
    const coords* c = img->pcoords;
    const int g = path == MDZG_PATH_GMP && c;
    mpfr_asprintf(&xmin, "%Ra", g ? c->xmin : img->xmin);      /* exact */
    mpfr_asprintf(&xmax, "%Ra", g ? c->xmax : img->xmax);
    mpfr_asprintf(&ymax, "%Ra", g ? c->ymax : img->ymax);
    mpfr_asprintf(&width, "%Ra", g ? c->width : img->width);
    if (img->family == FAMILY_JULIA) {
        mpfr_asprintf(&jre, "%Ra", img->u.julia.c_re);
        mpfr_asprintf(&jim, "%Ra", img->u.julia.c_im);
    }
    mdzg_view v = { img->fractal, img->family, path, img->real_width, img->real_height,
                    (long)img->precision, (long long)img->depth, xmin, xmax, ymax, width,
                    jre, jim };
    int rc = (xmin && xmax && ymax && width)
                 ? mdzg_frame_open(gpu_ctx, &v, (int32_t*)img->raw_data, &gpu_frame)
                 : MDZG_ERR_NOMEM;
    if (rc != MDZG_OK)
        gpu_frame = NULL;
    mpfr_free_str(xmin);
    mpfr_free_str(xmax);
    mpfr_free_str(ymax);
    mpfr_free_str(width);
    if (jre)
        mpfr_free_str(jre);
    if (jim)
        mpfr_free_str(jim);
    return rc;
}

static int gpu_line(image_info* img, int line, int path)
{
    if (rth_render_should_stop((rthdata*)img->rth_ptr)) {      /* RT_STOP */
        pthread_mutex_lock(&gpu_lock);
        if (gpu_frame) {
            mdzg_frame_stop(gpu_frame);   /* wakes rows waiting on the device */
            gpu_rows_left = 0;            /* the next call opens a new frame  */
        }
        pthread_mutex_unlock(&gpu_lock);
        return 0;
    }
    pthread_mutex_lock(&gpu_lock);
    if (!gpu_ctx && mdzg_create(0, &gpu_ctx) != MDZG_OK) {
        pthread_mutex_unlock(&gpu_lock);
        fprintf(stderr, "fractal_gpu: no MI355X device\n");
        return 0;                         /* nothing rendered, rows keep the sentinel */
    }
    if (!gpu_frame || gpu_rows_left == 0) {   /* the first row of a new render */
        if (gpu_frame)
            mdzg_frame_close(gpu_frame, NULL);
        gpu_frame = NULL;
        if (open_frame(img, path) != MDZG_OK) {
            pthread_mutex_unlock(&gpu_lock);
            return 0;
        }
        gpu_rows_left = img->real_height;
    }
    gpu_rows_left--;
    mdzg_frame* f = gpu_frame;
    pthread_mutex_unlock(&gpu_lock);
    const int ok = mdzg_frame_calculate_line(f, line);
    if (!ok) {                  /* stopped or failed: this render ends here */
        pthread_mutex_lock(&gpu_lock);
        if (gpu_frame == f)
            gpu_rows_left = 0;
        pthread_mutex_unlock(&gpu_lock);
    }
    return ok;
}

int fractal_gpu_mpfr_calculate_line(image_info* img, int line)
{
    return gpu_line(img, line, MDZG_PATH_MPFR);
}

int fractal_gpu_ld_calculate_line(image_info* img, int line)
{
    return gpu_line(img, line, MDZG_PATH_LD);
}

int fractal_gpu_gmp_calculate_line(image_info* img, int line)
{
    return gpu_line(img, line, MDZG_PATH_GMP);
}

/* the device context outlives renders; MDZ calls this at exit if it likes */
void fractal_gpu_shutdown(void)
{
    pthread_mutex_lock(&gpu_lock);
    if (gpu_frame)
        mdzg_frame_close(gpu_frame, NULL);
    gpu_frame = NULL;
    if (gpu_ctx)
        mdzg_destroy(gpu_ctx);
    gpu_ctx = NULL;
    pthread_mutex_unlock(&gpu_lock);
}
