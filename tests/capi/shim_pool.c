/*
 * shim_pool.c -- drive mdz_amd/shim/fractal_gpu.c the way MDZ's render
 * threads drive a line worker (src/render_threads.c:342-393): N worker
 * threads claim aa consecutive rows at a time from a shared counter under a
 * mutex, call the line callback for each, and leave on a 0 return or once
 * every row is claimed.  RT_STOP is modelled by a flag the pool raises once a
 * given number of claims were made; the shim reads it through
 * rth_render_should_stop, as MDZ's line workers do.
 *
 *   shim_pool PATH FRACTAL W H PRECISION DEPTH XMIN XMAX YMAX WIDTH
 *             THREADS AA STOP_AFTER OUT
 *
 * Two renders of the same image_info in one process: the first runs to the
 * end, the second raises the stop after STOP_AFTER claims.  raw_data starts
 * at the UINT_MAX sentinel each time (src/image_info.c:214-215).  Writes the
 * two maps (int32, H x W each) to OUT and prints the rows each render
 * completed.
 */
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <mdzg.h>

#include "fractal.h"
#include "render_threads.h"

struct rthdata {
    pthread_mutex_t lock;
    int stop;
    int next_line, claims, stop_after, aa, rows_done;
    image_info* img;
    int (*cb)(image_info*, int);
};

int rth_render_should_stop(rthdata* r)
{
    pthread_mutex_lock(&r->lock);
    const int s = r->stop;
    pthread_mutex_unlock(&r->lock);
    return s;
}

static void* worker(void* p)
{
    rthdata* r = (rthdata*)p;
    const int H = r->img->real_height;
    for (;;) {
        pthread_mutex_lock(&r->lock);
        if (r->stop || r->next_line >= H) {
            pthread_mutex_unlock(&r->lock);
            return NULL;
        }
        const int line = r->next_line;
        r->next_line += r->aa;
        r->claims++;
        if (r->stop_after >= 0 && r->claims > r->stop_after)
            r->stop = 1;
        pthread_mutex_unlock(&r->lock);
        for (int i = 0; i < r->aa && line + i < H; ++i) {
            if (!r->cb(r->img, line + i))
                return NULL;
            pthread_mutex_lock(&r->lock);
            r->rows_done++;
            pthread_mutex_unlock(&r->lock);
        }
    }
}

static int render(image_info* img, int (*cb)(image_info*, int), int threads, int aa,
                  int stop_after)
{
    rthdata r;
    memset(&r, 0, sizeof r);
    pthread_mutex_init(&r.lock, NULL);
    r.aa = aa;
    r.stop_after = stop_after;
    r.img = img;
    r.cb = cb;
    img->rth_ptr = &r;
    const size_t n = (size_t)img->real_width * (size_t)img->real_height;
    for (size_t k = 0; k < n; ++k)
        img->raw_data[k] = -1;                 /* UINT_MAX */
    pthread_t th[64];
    if (threads > 64)
        threads = 64;
    for (int t = 0; t < threads; ++t)
        pthread_create(&th[t], NULL, worker, &r);
    for (int t = 0; t < threads; ++t)
        pthread_join(th[t], NULL);
    pthread_mutex_destroy(&r.lock);
    return r.rows_done;
}

int main(int argc, char** argv)
{
    if (argc != 15) {
        fprintf(stderr, "usage: shim_pool PATH FRACTAL W H PRECISION DEPTH XMIN XMAX YMAX WIDTH "
                        "THREADS AA STOP_AFTER OUT\n");
        return 2;
    }
    const int path = atoi(argv[1]);
    image_info img;
    memset(&img, 0, sizeof img);
    img.fractal = atoi(argv[2]);
    img.family = FAMILY_MANDEL;
    img.real_width = atoi(argv[3]);
    img.real_height = atoi(argv[4]);
    img.precision = atol(argv[5]);
    img.depth = atoll(argv[6]);
    const mpfr_prec_t hp = img.precision > 80 ? img.precision : 80;
    mpfr_t* f[4] = { &img.xmin, &img.xmax, &img.ymax, &img.width };
    for (int i = 0; i < 4; ++i) {
        mpfr_init2(*f[i], hp);
        if (mpfr_set_str(*f[i], argv[7 + i], 0, MPFR_RNDN) != 0) {
            fprintf(stderr, "bad coordinate %s\n", argv[7 + i]);
            return 2;
        }
    }
    const int threads = atoi(argv[11]), aa = atoi(argv[12]), stop_after = atoi(argv[13]);
    const size_t n = (size_t)img.real_width * (size_t)img.real_height;
    img.raw_data = (int*)malloc(n * sizeof(int));
    int* first = (int*)malloc(n * sizeof(int));
    if (!img.raw_data || !first)
        return 3;
    int (*cb)(image_info*, int) = path == MDZG_PATH_LD ? fractal_gpu_ld_calculate_line
                                  : path == MDZG_PATH_GMP ? fractal_gpu_gmp_calculate_line
                                                          : fractal_gpu_mpfr_calculate_line;
    const int done1 = render(&img, cb, threads, aa, -1);
    memcpy(first, img.raw_data, n * sizeof(int));
    const int done2 = render(&img, cb, threads, aa, stop_after);
    FILE* out = fopen(argv[14], "wb");
    if (!out || fwrite(first, sizeof(int), n, out) != n ||
        fwrite(img.raw_data, sizeof(int), n, out) != n || fclose(out) != 0)
        return 4;
    printf("done1 %d done2 %d\n", done1, done2);
    fractal_gpu_shutdown();
    for (int i = 0; i < 4; ++i)
        mpfr_clear(*f[i]);
    free(img.raw_data);
    free(first);
    return 0;
}
