/*
 * frame.c -- the line-callback adapter (drop-in for next_line_cb).
 *
 * MDZ's worker threads call next_line_cb(img, line) once per row with
 * distinct rows, concurrently (src/render_threads.c:360-393); the callback
 * fills img->raw_data[line*W .. +W) and returns 1, or 0 once the render was
 * stopped (src/fractal.h:31-44, src/fractal.c:113-116, :236-240).
 *
 * Here the frame is cut into row bands of at least the context's band size
 * (MDZG_OPT_BAND_PIXELS, 2^18 pixels unless tuned) in pixels;
 * bands are submitted to the device in row order, a few at a time, as the
 * workers ask for their rows.  A call for row `line` makes sure its band (and
 * the next ones) are in flight, waits for its band, copies the row into the
 * caller's raw buffer and returns 1.  mdzg_frame_stop raises the device
 * cancel flag of the bands in flight; every later call returns 0 and rows
 * that were never rendered keep the caller's sentinel.
 */
#include <stdlib.h>
#include <string.h>

#include "mdzg_host.h"

struct mdzg_frame {
    mdzg_ctx* ctx;
    mdzg_view view;
    char* strs[6];
    int32_t* raw;
    int band_rows, nbands, next_submit, stopped, err;
    int inflight;                 /* bands kept in flight (MDZG_OPT_BAND_INFLIGHT) */
    mdzg_job** jobs;
    uint32_t** bufs;
    char* band_done;
    pthread_mutex_t* band_lock;   /* one waiter per band at a time */
    pthread_mutex_t lock;         /* submission, stop flag, stats */
    mdzg_stats sum;
};

static char* dup_or_null(const char* s)
{
    if (!s)
        return NULL;
    size_t n = strlen(s) + 1;
    char* d = (char*)malloc(n);
    if (d)
        memcpy(d, s, n);
    return d;
}

int mdzg_frame_open(mdzg_ctx* ctx, const mdzg_view* view, int32_t* raw, mdzg_frame** out)
{
    if (!ctx || !view || !raw || !out || view->real_width < 1 || view->real_height < 1)
        return MDZG_ERR_ARG;
    /* validate the view once up front */
    mdzg_pview pv;
    int rc = mdzg_view_parse(view, &pv);
    if (rc != MDZG_OK)
        return rc;
    mdzg_view_clear(&pv);

    mdzg_frame* f = (mdzg_frame*)calloc(1, sizeof *f);
    if (!f)
        return MDZG_ERR_NOMEM;
    pthread_mutex_init(&f->lock, NULL);
    f->ctx = ctx;
    f->view = *view;
    const char* src[6] = { view->xmin, view->xmax, view->ymax, view->width, view->julia_re,
                           view->julia_im };
    for (int i = 0; i < 6; ++i)
        f->strs[i] = dup_or_null(src[i]);
    f->view.xmin = f->strs[0];
    f->view.xmax = f->strs[1];
    f->view.ymax = f->strs[2];
    f->view.width = f->strs[3];
    f->view.julia_re = f->strs[4];
    f->view.julia_im = f->strs[5];
    f->raw = raw;
    const int W = view->real_width, H = view->real_height;
    f->band_rows = (int)((ctx->band_pixels + W - 1) / W);
    if (f->band_rows < 1)
        f->band_rows = 1;
    f->inflight = ctx->band_inflight;
    if (f->band_rows > H)
        f->band_rows = H;
    f->nbands = (H + f->band_rows - 1) / f->band_rows;
    f->jobs = (mdzg_job**)calloc((size_t)f->nbands, sizeof(mdzg_job*));
    f->bufs = (uint32_t**)calloc((size_t)f->nbands, sizeof(uint32_t*));
    f->band_done = (char*)calloc((size_t)f->nbands, 1);
    f->band_lock = (pthread_mutex_t*)calloc((size_t)f->nbands, sizeof(pthread_mutex_t));
    if (!f->jobs || !f->bufs || !f->band_done || !f->band_lock) {
        free(f->band_lock);
        f->band_lock = NULL;
        mdzg_frame_close(f, NULL);
        return MDZG_ERR_NOMEM;
    }
    for (int b = 0; b < f->nbands; ++b)
        pthread_mutex_init(&f->band_lock[b], NULL);
    *out = f;
    return MDZG_OK;
}

/* with f->lock held */
static int submit_upto(mdzg_frame* f, int last)
{
    while (f->next_submit <= last && f->next_submit < f->nbands) {
        const int b = f->next_submit;
        const int line0 = b * f->band_rows;
        int nl = f->band_rows;
        if (line0 + nl > f->view.real_height)
            nl = f->view.real_height - line0;
        f->bufs[b] = (uint32_t*)malloc((size_t)nl * (size_t)f->view.real_width * 4);
        if (!f->bufs[b])
            return MDZG_ERR_NOMEM;
        mdzg_tile t;
        t.view = f->view;
        t.line0 = line0;
        t.nlines = nl;
        t.col0 = 0;
        t.ncols = f->view.real_width;
        t.counts = f->bufs[b];
        int rc = mdzg_submit(f->ctx, &t, 1, &f->jobs[b]);
        if (rc != MDZG_OK)
            return rc;
        f->next_submit++;
    }
    return MDZG_OK;
}

static void add_stats(mdzg_stats* s, const mdzg_stats* a)
{
    s->kernel_ms += a->kernel_ms;
    s->wall_ms += a->wall_ms;
    s->pixel_iters += a->pixel_iters;
    s->launches += a->launches;
}

int mdzg_frame_calculate_line(mdzg_frame* f, int line)
{
    if (!f || line < 0 || line >= f->view.real_height)
        return 0;
    const int b = line / f->band_rows;
    pthread_mutex_lock(&f->lock);
    if (f->stopped || f->err) {
        pthread_mutex_unlock(&f->lock);
        return 0;
    }
    int rc = submit_upto(f, b + f->inflight - 1);
    if (rc != MDZG_OK) {
        f->err = rc;
        pthread_mutex_unlock(&f->lock);
        return 0;
    }
    mdzg_job* job = f->jobs[b];
    pthread_mutex_unlock(&f->lock);

    /* wait outside the frame lock so mdzg_frame_stop can cancel meanwhile */
    pthread_mutex_lock(&f->band_lock[b]);
    if (!f->band_done[b]) {
        mdzg_stats st;
        memset(&st, 0, sizeof st);
        rc = mdzg_wait(job, &st);
        if (rc == MDZG_OK) {
            f->band_done[b] = 1;
            pthread_mutex_lock(&f->lock);
            add_stats(&f->sum, &st);
            pthread_mutex_unlock(&f->lock);
        }
    }
    pthread_mutex_unlock(&f->band_lock[b]);
    if (rc != MDZG_OK) {
        pthread_mutex_lock(&f->lock);
        if (rc != MDZG_ERR_STOPPED)
            f->err = rc;
        pthread_mutex_unlock(&f->lock);
        return 0;
    }
    const int W = f->view.real_width;
    memcpy(f->raw + (size_t)line * W,
           f->bufs[b] + (size_t)(line - b * f->band_rows) * W, (size_t)W * 4);
    return 1;
}

void mdzg_frame_stop(mdzg_frame* f)
{
    if (!f)
        return;
    pthread_mutex_lock(&f->lock);
    f->stopped = 1;
    for (int b = 0; b < f->next_submit; ++b)
        if (f->jobs[b] && !f->band_done[b])
            mdzg_cancel(f->jobs[b]);
    pthread_mutex_unlock(&f->lock);
}

int mdzg_frame_close(mdzg_frame* f, mdzg_stats* st)
{
    if (!f)
        return MDZG_ERR_ARG;
    int rc = f->err;
    if (f->jobs)
        for (int b = 0; b < f->nbands; ++b)
            if (f->jobs[b])
                mdzg_job_free(f->jobs[b]);
    if (f->bufs)
        for (int b = 0; b < f->nbands; ++b)
            free(f->bufs[b]);
    if (st)
        *st = f->sum;
    for (int i = 0; i < 6; ++i)
        free(f->strs[i]);
    free(f->jobs);
    free(f->bufs);
    free(f->band_done);
    if (f->band_lock) {
        for (int b = 0; b < f->nbands; ++b)
            pthread_mutex_destroy(&f->band_lock[b]);
        free(f->band_lock);
    }
    pthread_mutex_destroy(&f->lock);
    rc = rc ? rc : (f->stopped ? MDZG_ERR_STOPPED : MDZG_OK);
    free(f);
    return rc;
}
