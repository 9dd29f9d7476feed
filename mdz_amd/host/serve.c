/*
 * serve.c -- continuous-batching tile server (C99).
 *
 * MDZ has no serving form (TODO:15-16 asks for a "rendering daemon" with
 * queued jobs); this is the device-side one.  Clients submit tiles (views at
 * any depth and precision) at any time; a scheduler thread runs the device in
 * slices:
 *
 *   admit   pending requests, oldest first, while the HBM budget (a first-fit
 *           arena carved out of one allocation) and the per-class budget of
 *           resident pixels (a multiple of the lanes the class's slice kernel
 *           keeps resident: the VGPR-occupancy budget) allow;
 *   stage   the new requests' start values (host MPFR / x87 chains), count
 *           sentinels and zeroed iteration counters;
 *   slice   one launch per arithmetic class over every unfinished pixel of
 *           every active request: each pixel resumes from its saved orbit
 *           state (z, z^2, iterations done), runs at most `slice_iters` more
 *           iterations, and either finishes (count written, the request's
 *           remaining counter decremented) or saves its state again;
 *   retire  requests whose remaining counter reached zero: counts copied to
 *           the client, latency recorded, memory returned.
 *
 * Requests that arrive while others are on the device join at the next slice,
 * so their work overlaps on the device with everything already running; the
 * orbit state stays resident in HBM between slices.  A pixel's iterations are
 * the same whether its request runs alone or in a mix (the arithmetic is
 * exact and per pixel), which the tests check.
 *
 * MDZG_SERVE_PER_REQUEST is the baseline form: every request is submitted on
 * arrival as its own asynchronous job (one launch per tile).
 */
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mdzg_host.h"

#define SENT 0xffffffffu
#define ALIGN_UP(x, a) (((x) + (size_t)(a) - 1) & ~((size_t)(a) - 1))
#define MAX_KID MDZG_NKERNELS

typedef struct sreq {
    uint64_t id;
    mdzg_tile tile;
    mdzg_pview pv;
    int parsed;
    uint32_t npix;
    size_t bytes, off;
    int slot;
    int status;                 /* 0 pending, 1 active, 2 done, < 0 error */
    double t_submit, t_admit, t_done;
    double pixel_iters;
    double work_left;           /* unfinished pixels x depth: the slice queue's order */
    mdzg_job* job;              /* per-request mode */
    /* device addresses (batched mode) */
    uint32_t *d_xs, *d_ys, *d_jc, *d_out, *d_state;
    int32_t* d_wzs;
    double xmin, ymax, width, jre, jim;
    /* image requests: counts coloured on the device when the request is done */
    int image;
    int want_counts;
    mdzg_palette pal;
    mdzg_image_req ir;          /* per-request mode */
    uint32_t *d_rgb, *d_pal;
    unsigned long long* d_iters;
} sreq;

typedef struct ablock {
    size_t off, size;
    int used;
} ablock;

struct mdzg_server {
    mdzg_ctx* ctx;
    mdzg_server_cfg cfg;
    pthread_t thr;
    int thr_started;
    pthread_mutex_t lock;
    pthread_cond_t cv_work, cv_done;
    int shutdown;

    sreq** reqs;
    size_t nreq, capreq;
    size_t* pend;               /* FIFO of request indices */
    size_t pend_head, pend_tail, cappend;
    sreq** act;
    int nact, capact;

    /* device memory */
    char* d_arena;
    size_t arena_cap, arena_used, arena_peak;
    ablock* blocks;
    int nblocks, capblocks;
    unsigned int* d_remaining;
    unsigned int* h_remaining;
    int* slot_used;
    int nslots;
    mdzg_dtile* d_desc;
    mdzg_dtile* h_desc;
    int capdesc;
    unsigned long long* d_ctr;
    void* h_stage;
    size_t capstage;
    hipStream_t streams[MAX_KID];
    hipEvent_t ev_fork, ev_join[MAX_KID];

    int occ_slice[MAX_KID];
    long long budget_pix[MAX_KID];
    long long act_pix[MAX_KID];
    int slice_iters[MAX_KID];
    /* tests: the id of a request whose staging fails as a device error would
       (environment variable MDZG_FAILPOINT_STAGE read at mdzg_server_create;
       0 = none) */
    uint64_t fail_stage_id;

    mdzg_serve_stats st;
};

/* ---- arena (first fit, coalescing) -------------------------------------- */
static int arena_alloc(mdzg_server* s, size_t bytes, size_t* off)
{
    bytes = ALIGN_UP(bytes, 256);
    for (int i = 0; i < s->nblocks; ++i) {
        ablock* b = &s->blocks[i];
        if (b->used || b->size < bytes)
            continue;
        if (b->size > bytes) {
            if (s->nblocks == s->capblocks) {
                int nc = s->capblocks * 2;
                ablock* nb = (ablock*)realloc(s->blocks, sizeof(ablock) * (size_t)nc);
                if (!nb)
                    return -1;
                s->blocks = nb;
                s->capblocks = nc;
                b = &s->blocks[i];
            }
            memmove(&s->blocks[i + 2], &s->blocks[i + 1],
                    sizeof(ablock) * (size_t)(s->nblocks - i - 1));
            s->blocks[i + 1].off = b->off + bytes;
            s->blocks[i + 1].size = b->size - bytes;
            s->blocks[i + 1].used = 0;
            s->nblocks++;
            b->size = bytes;
        }
        b->used = 1;
        *off = b->off;
        s->arena_used += bytes;
        if (s->arena_used > s->arena_peak)
            s->arena_peak = s->arena_used;
        return 0;
    }
    return -1;
}

static void arena_free(mdzg_server* s, size_t off)
{
    for (int i = 0; i < s->nblocks; ++i) {
        if (s->blocks[i].off != off || !s->blocks[i].used)
            continue;
        s->blocks[i].used = 0;
        s->arena_used -= s->blocks[i].size;
        if (i + 1 < s->nblocks && !s->blocks[i + 1].used) {
            s->blocks[i].size += s->blocks[i + 1].size;
            memmove(&s->blocks[i + 1], &s->blocks[i + 2],
                    sizeof(ablock) * (size_t)(s->nblocks - i - 2));
            s->nblocks--;
        }
        if (i > 0 && !s->blocks[i - 1].used) {
            s->blocks[i - 1].size += s->blocks[i].size;
            memmove(&s->blocks[i], &s->blocks[i + 1],
                    sizeof(ablock) * (size_t)(s->nblocks - i - 1));
            s->nblocks--;
        }
        return;
    }
}

/* ---- request layout -------------------------------------------------------- */
static size_t req_bytes(const sreq* r)
{
    const size_t npix = r->npix;
    size_t b = 0;
    if (r->pv.vs) {
        const size_t vsb = 4 * (size_t)r->pv.vs;
        b += ALIGN_UP(((size_t)r->tile.ncols + (size_t)r->tile.nlines + 2) * vsb, 256);
        b += ALIGN_UP(npix * 4 * vsb, 256);              /* state: 4 values */
    } else {
        b += ALIGN_UP(npix * 32, 256);                   /* state: 4 doubles */
    }
    b += ALIGN_UP(npix * 4, 256);                        /* counts */
    b += ALIGN_UP(npix * 4, 256);                        /* iterations done */
    if (r->image)                                        /* colours, palette, iteration sum */
        b += ALIGN_UP((size_t)r->ir.user_w * (size_t)r->ir.user_h * 4, 256) + 1024 + 256;
    return b;
}

static int default_slice_iters(int nw)
{
    if (nw == 0) return 8192;
    if (nw <= 2) return 2048;
    if (nw <= 4) return 1024;
    if (nw <= 8) return 512;
    if (nw <= 16) return 128;
    return 32;
}

static void fill_info(const sreq* r, mdzg_req_info* info)
{
    info->status = r->status;
    info->submit_ms = r->t_submit;
    info->admit_ms = r->t_admit;
    info->done_ms = r->t_done;
    info->latency_ms = r->status == 2 ? r->t_done - r->t_submit : -1.0;
    info->pixel_iters = r->pixel_iters;
}

/* pixel-iterations of n counts (a zero count stands for the depth), summed
   in integers as mdzg_take_counts does */
static double count_iters(const uint32_t* c, size_t n, double depth)
{
    uint64_t sum = 0, zeros = 0;
    for (size_t k = 0; k < n; ++k) {
        sum += c[k];
        zeros += (c[k] == 0);
    }
    return (double)sum + (double)zeros * depth;
}

/* ---- batched mode: one slice ------------------------------------------------ */
/* Stage the newly admitted requests.  A request whose staging fails gets
   the error as its status and takes no part in the slices; the others go
   on.  The return value is an error only when the staging buffer cannot be
   had, which fails every new request. */
static int stage_one(mdzg_server* s, sreq* r, size_t* hoff, hipStream_t st);

static int stage_new(mdzg_server* s, sreq** nw, int nnew, hipStream_t st)
{
    /* start values go through the pinned staging buffer */
    size_t need = 0;
    for (int i = 0; i < nnew; ++i) {
        if (nw[i]->pv.vs)
            need += ALIGN_UP(((size_t)nw[i]->tile.ncols + (size_t)nw[i]->tile.nlines + 2) *
                             4 * (size_t)nw[i]->pv.vs, 256);
        if (nw[i]->image)
            need += 1024;
    }
    if (need > s->capstage) {
        if (s->h_stage) {
            hipStreamSynchronize(st);
            hipHostFree(s->h_stage);
        }
        s->capstage = ALIGN_UP(need * 2, 1 << 20);
        if (hipHostMalloc(&s->h_stage, s->capstage, hipHostMallocDefault) != hipSuccess) {
            s->h_stage = NULL;
            s->capstage = 0;
            return MDZG_ERR_NOMEM;
        }
    }
    size_t hoff = 0;
    for (int i = 0; i < nnew; ++i) {
        const int rc = stage_one(s, nw[i], &hoff, st);
        if (rc != MDZG_OK)
            nw[i]->status = rc;
        if (nw[i]->parsed) {
            mdzg_view_clear(&nw[i]->pv);
            nw[i]->parsed = 0;
        }
    }
    return MDZG_OK;
}

static int stage_one(mdzg_server* s, sreq* r, size_t* hoffp, hipStream_t st)
{
    size_t hoff = *hoffp;
    {
        const mdzg_tile* t = &r->tile;
        if (r->id == s->fail_stage_id)
            return MDZG_ERR_DEVICE;
        char* base = s->d_arena + r->off;
        size_t o = 0;
        if (r->pv.vs) {
            const size_t vsb = 4 * (size_t)r->pv.vs;
            const size_t cb = ((size_t)t->ncols + (size_t)t->nlines + 2) * vsb;
            char* H = (char*)s->h_stage + hoff;
            uint32_t* xs = (uint32_t*)H;
            uint32_t* ys = (uint32_t*)(H + vsb * (size_t)t->ncols);
            uint32_t* jc = (uint32_t*)(H + vsb * ((size_t)t->ncols + (size_t)t->nlines));
            int rc = mdzg_coords_fill(&r->pv, t->line0, t->nlines, t->col0, t->ncols, xs, ys, jc);
            if (rc != MDZG_OK) {
                *hoffp = hoff;
                return rc;
            }
            r->d_xs = (uint32_t*)base;
            r->d_ys = (uint32_t*)(base + vsb * (size_t)t->ncols);
            r->d_jc = (uint32_t*)(base + vsb * ((size_t)t->ncols + (size_t)t->nlines));
            if (hipMemcpyAsync(base, H, cb, hipMemcpyHostToDevice, st) != hipSuccess)
                return MDZG_ERR_DEVICE;
            hoff += ALIGN_UP(cb, 256);
            o += ALIGN_UP(cb, 256);
            r->d_state = (uint32_t*)(base + o);
            o += ALIGN_UP((size_t)r->npix * 4 * vsb, 256);
        } else {
            mdzg_coords_fp64(&r->pv, &r->xmin, &r->ymax, &r->width, &r->jre, &r->jim);
            r->d_state = (uint32_t*)(base + o);
            o += ALIGN_UP((size_t)r->npix * 32, 256);
        }
        r->d_out = (uint32_t*)(base + o);
        o += ALIGN_UP((size_t)r->npix * 4, 256);
        r->d_wzs = (int32_t*)(base + o);
        o += ALIGN_UP((size_t)r->npix * 4, 256);
        if (r->image) {
            r->d_rgb = (uint32_t*)(base + o);
            o += ALIGN_UP((size_t)r->ir.user_w * (size_t)r->ir.user_h * 4, 256);
            r->d_pal = (uint32_t*)(base + o);
            r->d_iters = (unsigned long long*)(base + o + 1024);
            char* H = (char*)s->h_stage + hoff;
            memcpy(H, r->pal.colours, 4 * (size_t)r->pal.n);
            if (hipMemcpyAsync(r->d_pal, H, 4 * (size_t)r->pal.n, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
                hipMemsetD32Async((hipDeviceptr_t)r->d_iters, 0, 2, st) != hipSuccess)
                return MDZG_ERR_DEVICE;
            hoff += 1024;
        }
        if (hipMemsetD32Async((hipDeviceptr_t)r->d_out, (int)SENT, r->npix, st) != hipSuccess ||
            hipMemsetD32Async((hipDeviceptr_t)r->d_wzs, 0, r->npix, st) != hipSuccess ||
            hipMemsetD32Async((hipDeviceptr_t)(s->d_remaining + r->slot), (int)r->npix, 1, st) !=
                hipSuccess) {
            *hoffp = hoff;
            return MDZG_ERR_DEVICE;
        }
    }
    *hoffp = hoff;
    return MDZG_OK;
}

static int run_slice(mdzg_server* s)
{
    mdzg_ctx* c = s->ctx;
    hipStream_t st = s->streams[0];
    /* descriptors, grouped by class, oldest request first */
    if (s->nact > s->capdesc) {
        if (s->d_desc) hipFree(s->d_desc);
        if (s->h_desc) hipHostFree(s->h_desc);
        s->capdesc = s->nact * 2;
        if (hipMalloc((void**)&s->d_desc, sizeof(mdzg_dtile) * (size_t)s->capdesc) != hipSuccess ||
            hipHostMalloc((void**)&s->h_desc, sizeof(mdzg_dtile) * (size_t)s->capdesc,
                          hipHostMallocDefault) != hipSuccess)
            return MDZG_ERR_NOMEM;
    }
    int first[MAX_KID], cnt[MAX_KID];
    uint64_t total[MAX_KID];
    int n = 0;
    /* Each class's requests go in the slice queue with the least work left
       first (oldest first among equals).  A class whose requests are alike
       gives every pixel slice_iters iterations a slice.  A class whose
       requests differ in work by more than 8x caps each lane at
       2 x slice_iters iterations instead and lets a pixel run on within
       that: the lanes claim from the front, so a small request advances as
       it would alone instead of sharing every slice with a large one, and
       the large one takes the rest of the lanes. */
    sreq** order = (sreq**)malloc(sizeof(sreq*) * (size_t)(s->nact > 0 ? s->nact : 1));
    if (!order)
        return MDZG_ERR_NOMEM;
    int lane_cap[MAX_KID];
    for (int k = 0; k < MAX_KID; ++k) {
        first[k] = n;
        cnt[k] = 0;
        total[k] = 0;
        lane_cap[k] = 0;
        int m = 0;
        for (int i = 0; i < s->nact; ++i) {
            sreq* r = s->act[i];
            if (r->pv.kid != k || r->status < 0)
                continue;
            int j = m++;
            while (j > 0 && (order[j - 1]->work_left > r->work_left ||
                             (order[j - 1]->work_left == r->work_left && order[j - 1]->id > r->id))) {
                order[j] = order[j - 1];
                --j;
            }
            order[j] = r;
        }
        if (m >= 2 && order[m - 1]->work_left > 8.0 * order[0]->work_left)
            lane_cap[k] = 1;
        for (int oi = 0; oi < m; ++oi) {
            sreq* r = order[oi];
            mdzg_dtile* d = &s->h_desc[n++];
            memset(d, 0, sizeof *d);
            d->out = r->d_out;
            d->xs = r->d_xs;
            d->ys = r->d_ys;
            d->jc = r->d_jc;
            d->xmin = r->xmin;
            d->ymax = r->ymax;
            d->width = r->width;
            d->jre = r->jre;
            d->jim = r->jim;
            d->pix_begin = total[k];
            d->ncols = r->tile.ncols;
            d->nrows = r->tile.nlines;
            d->col0 = r->tile.col0;
            d->line0 = r->tile.line0;
            d->W = r->tile.view.real_width;
            d->fractal = r->tile.view.fractal;
            d->family = r->tile.view.family;
            d->kclr = r->pv.kclr;
            d->depth = (int32_t)r->tile.view.depth;
            d->vs = r->pv.vs;
            d->gprec = r->tile.view.path == MDZG_PATH_GMP ? r->pv.nw : 0;
            d->state = r->d_state;
            d->wzs = r->d_wzs;
            d->remaining = s->d_remaining + r->slot;
            total[k] += r->npix;
            cnt[k]++;
        }
    }
    free(order);
    if (n == 0)
        return MDZG_OK;          /* every active request failed its staging */
    if (hipMemcpyAsync(s->d_desc, s->h_desc, sizeof(mdzg_dtile) * (size_t)n,
                       hipMemcpyHostToDevice, st) != hipSuccess)
        return MDZG_ERR_DEVICE;
    /* the slice: the queue counters reset by a fill kernel, then one slice
       launch per class, concurrently -- one graph replay, or launched one by
       one with a fork and join over the server's streams */
    mdzg_gnode reset, par[MAX_KID];
    int nk = 0;
    if (mdzg_dev_node_fill(&reset, (uint32_t*)s->d_ctr, 2 * MAX_KID, 0u) != 0)
        return MDZG_ERR_DEVICE;
    for (int k = 0; k < MAX_KID; ++k) {
        if (!cnt[k])
            continue;
        const uint64_t cap = (uint64_t)c->ncu * (uint64_t)s->occ_slice[k];
        const uint64_t want = (total[k] + 255) / 256;
        int grid = (int)(want < cap ? want : cap);
        if (grid < 1) grid = 1;
        const int it = s->slice_iters[k];
        if (mdzg_dev_node_slice(&par[nk++], k, s->d_desc + first[k], cnt[k], total[k],
                                s->d_ctr + k, grid, lane_cap[k] ? INT32_MAX : it,
                                lane_cap[k] ? 2 * it : INT32_MAX) != 0)
            return MDZG_ERR_DEVICE;
    }
    if (c->graphs && nk >= 2) {
        if (mdzg_graph_run(c->gc, &reset, par, nk, NULL, st) != 0)
            return MDZG_ERR_DEVICE;
    } else {
        if (mdzg_dev_node_launch(&reset, st) != 0)
            return MDZG_ERR_DEVICE;
        hipEventRecord(s->ev_fork, st);
        for (int q = 0; q < nk; ++q) {
            hipStream_t ks = s->streams[q];
            if (q > 0)
                hipStreamWaitEvent(ks, s->ev_fork, 0);
            if (mdzg_dev_node_launch(&par[q], ks) != 0)
                return MDZG_ERR_DEVICE;
            if (q > 0) {
                hipEventRecord(s->ev_join[q], ks);
                hipStreamWaitEvent(st, s->ev_join[q], 0);
            }
        }
    }
    if (hipMemcpyAsync(s->h_remaining, s->d_remaining, sizeof(unsigned int) * (size_t)s->nslots,
                       hipMemcpyDeviceToHost, st) != hipSuccess)
        return MDZG_ERR_DEVICE;
    if (hipStreamSynchronize(st) != hipSuccess)
        return MDZG_ERR_DEVICE;
    s->st.slices++;
    return MDZG_OK;
}

/* a finished request's counts to the client */
static int retire_counts(sreq* r)
{
    if (hipMemcpy(r->tile.counts, r->d_out, (size_t)r->npix * 4, hipMemcpyDeviceToHost) !=
        hipSuccess)
        return MDZG_ERR_DEVICE;
    r->pixel_iters = count_iters(r->tile.counts, r->npix, (double)r->tile.view.depth);
    return MDZG_OK;
}

/* a finished image request: its counts coloured where they lie, the colours
   (and the counts, if asked for) to the client */
static int retire_image(mdzg_server* s, sreq* r)
{
    hipStream_t st = s->streams[0];
    unsigned long long iters = 0;
    const size_t nrgb = (size_t)r->ir.user_w * (size_t)r->ir.user_h;
    if (mdzg_dev_launch_colour(r->d_out, r->tile.ncols, r->ir.user_w, r->ir.user_h, r->ir.aa,
                               r->d_pal, r->pal.n, r->pal.offset, r->pal.colour_scale,
                               r->pal.interpolate, (uint32_t)r->tile.view.depth, r->d_rgb,
                               r->d_iters, MDZG_COLOUR_BLOCKS_PER_CU * s->ctx->ncu, st) != 0 ||
        hipMemcpyAsync(r->ir.rgb, r->d_rgb, nrgb * 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipMemcpyAsync(&iters, r->d_iters, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        (r->want_counts && hipMemcpyAsync(r->tile.counts, r->d_out, (size_t)r->npix * 4,
                                          hipMemcpyDeviceToHost, st) != hipSuccess) ||
        hipStreamSynchronize(st) != hipSuccess)
        return MDZG_ERR_DEVICE;
    r->pixel_iters = (double)iters;
    return MDZG_OK;
}

static void* sched_batched(void* arg)
{
    mdzg_server* s = (mdzg_server*)arg;
    hipSetDevice(s->ctx->device);
    sreq** fresh = NULL;
    int capfresh = 0;
    pthread_mutex_lock(&s->lock);
    for (;;) {
        while (!s->shutdown && s->nact == 0 && s->pend_head == s->pend_tail)
            pthread_cond_wait(&s->cv_work, &s->lock);
        if (s->shutdown && s->nact == 0 && s->pend_head == s->pend_tail)
            break;
        /* admission, oldest first.  A request that does not fit its class's
           budget of resident pixels holds back the later requests of its
           class (first come, first served within a class) but not those of
           other classes: a 1024-bit tile waiting for room does not keep an
           FP64 tile off the device.  No free slot or no arena block stops
           admission for this round. */
        int nnew = 0;
        int blocked[MAX_KID];
        memset(blocked, 0, sizeof blocked);
        size_t keep = s->pend_head;
        int full = 0;
        for (size_t pi = s->pend_head; pi < s->pend_tail; ++pi) {
            const size_t ri = s->pend[pi];
            sreq* r = s->reqs[ri];
            const int k = r->pv.kid;
            if (full || blocked[k] ||
                (s->act_pix[k] > 0 && s->act_pix[k] + r->npix > s->budget_pix[k])) {
                blocked[k] = 1;
                s->pend[keep++] = ri;
                continue;
            }
            int slot = -1;
            for (int i = 0; i < s->nslots; ++i)
                if (!s->slot_used[i]) {
                    slot = i;
                    break;
                }
            size_t off = 0;
            if (slot < 0 || arena_alloc(s, r->bytes, &off) != 0) {
                full = 1;
                s->pend[keep++] = ri;
                continue;
            }
            s->slot_used[slot] = 1;
            r->slot = slot;
            r->off = off;
            r->status = 1;
            r->work_left = (double)r->npix * (double)r->tile.view.depth;
            r->t_admit = mdzg_now_ms();
            s->act_pix[k] += r->npix;
            if (s->nact == s->capact) {
                s->capact = s->capact ? 2 * s->capact : 64;
                s->act = (sreq**)realloc(s->act, sizeof(sreq*) * (size_t)s->capact);
            }
            s->act[s->nact++] = r;
            if (nnew == capfresh) {
                capfresh = capfresh ? 2 * capfresh : 64;
                fresh = (sreq**)realloc(fresh, sizeof(sreq*) * (size_t)capfresh);
            }
            fresh[nnew++] = r;
            s->st.admitted++;
        }
        s->pend_tail = keep;
        pthread_mutex_unlock(&s->lock);

        int rc = MDZG_OK;
        if (nnew)
            rc = stage_new(s, fresh, nnew, s->streams[0]);
        if (rc != MDZG_OK)
            for (int i = 0; i < nnew; ++i)
                fresh[i]->status = rc;   /* no staging buffer: every new request fails */
        rc = MDZG_OK;
        if (s->nact)
            rc = run_slice(s);
        /* a failed request's memory goes back to the arena only after
           everything launched on the server's streams has finished with it */
        int failed = rc != MDZG_OK;
        for (int i = 0; i < s->nact && !failed; ++i)
            failed = s->act[i]->status < 0;
        if (failed)
            for (int k = 0; k < MAX_KID; ++k)
                if (s->streams[k])
                    hipStreamSynchronize(s->streams[k]);

        /* retire finished requests; a failed slice fails every active
           request, a failed staging only its own */
        pthread_mutex_lock(&s->lock);
        for (int i = 0; i < s->nact;) {
            sreq* r = s->act[i];
            const int done = (rc != MDZG_OK) || r->status < 0 || s->h_remaining[r->slot] == 0;
            if (!done) {
                r->work_left = (double)s->h_remaining[r->slot] * (double)r->tile.view.depth;
                ++i;
                continue;
            }
            if (r->status < 0) {
                /* its staging failed: the status already holds the error */
            } else if (rc == MDZG_OK) {
                pthread_mutex_unlock(&s->lock);
                const int rr = r->image ? retire_image(s, r) : retire_counts(r);
                pthread_mutex_lock(&s->lock);
                r->status = rr == MDZG_OK ? 2 : rr;
                s->st.completed += rr == MDZG_OK;
                s->st.pixel_iters += r->pixel_iters;
            } else {
                r->status = rc;
            }
            r->t_done = mdzg_now_ms();
            s->act_pix[r->pv.kid] -= r->npix;
            s->slot_used[r->slot] = 0;
            arena_free(s, r->off);
            s->act[i] = s->act[--s->nact];
        }
        /* keep the active list in admission order (oldest first) */
        for (int i = 1; i < s->nact; ++i)
            for (int j = i; j > 0 && s->act[j - 1]->id > s->act[j]->id; --j) {
                sreq* t = s->act[j];
                s->act[j] = s->act[j - 1];
                s->act[j - 1] = t;
            }
        pthread_cond_broadcast(&s->cv_done);
    }
    pthread_mutex_unlock(&s->lock);
    free(fresh);
    return NULL;
}

/* ---- per-request mode (baseline) ------------------------------------------- */
static void* sched_per_request(void* arg)
{
    mdzg_server* s = (mdzg_server*)arg;
    hipSetDevice(s->ctx->device);
    pthread_mutex_lock(&s->lock);
    for (;;) {
        while (!s->shutdown && s->nact == 0 && s->pend_head == s->pend_tail)
            pthread_cond_wait(&s->cv_work, &s->lock);
        if (s->shutdown && s->nact == 0 && s->pend_head == s->pend_tail)
            break;
        while (s->pend_head != s->pend_tail) {
            sreq* r = s->reqs[s->pend[s->pend_head++]];
            pthread_mutex_unlock(&s->lock);
            int rc = r->image ? mdzg_rt_submit_image(s->ctx, &r->tile, &r->ir, &r->job)
                              : mdzg_submit(s->ctx, &r->tile, 1, &r->job);
            pthread_mutex_lock(&s->lock);
            r->t_admit = mdzg_now_ms();
            s->st.admitted++;
            if (rc != MDZG_OK) {
                r->status = rc;
                r->t_done = r->t_admit;
                continue;
            }
            r->status = 1;
            if (s->nact == s->capact) {
                s->capact = s->capact ? 2 * s->capact : 64;
                s->act = (sreq**)realloc(s->act, sizeof(sreq*) * (size_t)s->capact);
            }
            s->act[s->nact++] = r;
        }
        int progressed = 0;
        for (int i = 0; i < s->nact;) {
            sreq* r = s->act[i];
            if (mdzg_poll(r->job) != 1) {
                ++i;
                continue;
            }
            pthread_mutex_unlock(&s->lock);
            mdzg_stats st;
            int rc = mdzg_wait(r->job, &st);
            mdzg_job_free(r->job);
            pthread_mutex_lock(&s->lock);
            r->job = NULL;
            r->pixel_iters = st.pixel_iters;
            r->status = rc == MDZG_OK ? 2 : rc;
            r->t_done = mdzg_now_ms();
            s->st.completed += rc == MDZG_OK;
            s->st.pixel_iters += st.pixel_iters;
            s->act[i] = s->act[--s->nact];
            progressed = 1;
        }
        if (progressed)
            pthread_cond_broadcast(&s->cv_done);
        else if (s->nact) {
            pthread_mutex_unlock(&s->lock);
            struct timespec ts = { 0, 50000 };
            nanosleep(&ts, NULL);
            pthread_mutex_lock(&s->lock);
        }
    }
    pthread_mutex_unlock(&s->lock);
    return NULL;
}

/* ---- public API ------------------------------------------------------------------ */
int mdzg_server_create(mdzg_ctx* ctx, const mdzg_server_cfg* cfg, mdzg_server** out)
{
    if (!ctx || !out)
        return MDZG_ERR_ARG;
    *out = NULL;
    mdzg_server* s = (mdzg_server*)calloc(1, sizeof *s);
    if (!s)
        return MDZG_ERR_NOMEM;
    s->ctx = ctx;
    if (cfg)
        s->cfg = *cfg;
    {
        const char* fp = getenv("MDZG_FAILPOINT_STAGE");
        s->fail_stage_id = (fp && *fp) ? strtoull(fp, NULL, 10) : 0;
    }
    if (s->cfg.mode != MDZG_SERVE_BATCHED && s->cfg.mode != MDZG_SERVE_PER_REQUEST) {
        free(s);
        return MDZG_ERR_ARG;
    }
    if (s->cfg.hbm_budget == 0)
        s->cfg.hbm_budget = (size_t)8 << 30;
    if (s->cfg.occupancy_budget <= 0.0)
        s->cfg.occupancy_budget = 4.0;
    if (s->cfg.max_requests <= 0)
        s->cfg.max_requests = 4096;
    pthread_mutex_init(&s->lock, NULL);
    pthread_cond_init(&s->cv_work, NULL);
    pthread_cond_init(&s->cv_done, NULL);
    hipSetDevice(ctx->device);
    int rc = MDZG_OK;
    if (s->cfg.mode == MDZG_SERVE_BATCHED) {
        s->arena_cap = s->cfg.hbm_budget;
        s->nslots = s->cfg.max_requests;
        s->capblocks = 64;
        s->blocks = (ablock*)calloc((size_t)s->capblocks, sizeof(ablock));
        s->slot_used = (int*)calloc((size_t)s->nslots, sizeof(int));
        if (!s->blocks || !s->slot_used) {
            rc = MDZG_ERR_NOMEM;
            goto fail;
        }
        s->blocks[0].off = 0;
        s->blocks[0].size = s->arena_cap;
        s->nblocks = 1;
        if (hipMalloc((void**)&s->d_arena, s->arena_cap) != hipSuccess ||
            hipMalloc((void**)&s->d_remaining, sizeof(unsigned int) * (size_t)s->nslots) != hipSuccess ||
            hipHostMalloc((void**)&s->h_remaining, sizeof(unsigned int) * (size_t)s->nslots,
                          hipHostMallocDefault) != hipSuccess ||
            hipMalloc((void**)&s->d_ctr, 16 * MAX_KID) != hipSuccess) {
            rc = MDZG_ERR_NOMEM;
            goto fail;
        }
        for (int k = 0; k < MAX_KID; ++k) {
            if (hipStreamCreateWithFlags(&s->streams[k], hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&s->ev_join[k], hipEventDisableTiming) != hipSuccess) {
                rc = MDZG_ERR_DEVICE;
                goto fail;
            }
            s->occ_slice[k] = mdzg_dev_occupancy_slice(k);
            if (s->occ_slice[k] < 1)
                s->occ_slice[k] = 1;
            const int lanes = ctx->ncu * s->occ_slice[k] * 256;
            s->budget_pix[k] = (long long)(s->cfg.occupancy_budget * (double)lanes);
            int nw = 0;
            if (k >= MDZG_KID_GMP0)       /* GMP path: limbs of 64 bits */
                nw = 2 * mdzg_gmp_limbs[k - MDZG_KID_GMP0];
            else if (k > 0)
                nw = mdzg_nwords[(k - 1) >> 1];
            s->slice_iters[k] = s->cfg.slice_iters > 0 ? s->cfg.slice_iters : default_slice_iters(nw);
        }
        if (hipEventCreateWithFlags(&s->ev_fork, hipEventDisableTiming) != hipSuccess) {
            rc = MDZG_ERR_DEVICE;
            goto fail;
        }
    }
    if (pthread_create(&s->thr, NULL,
                       s->cfg.mode == MDZG_SERVE_BATCHED ? sched_batched : sched_per_request, s)) {
        rc = MDZG_ERR_NOMEM;
        goto fail;
    }
    s->thr_started = 1;
    *out = s;
    return MDZG_OK;
fail:
    mdzg_server_destroy(s);
    return rc;
}

static int submit(mdzg_server* s, const mdzg_tile* t, int aa, const mdzg_palette* pal,
                  uint32_t* rgb, uint64_t* id)
{
    if (!s || !t || !id || t->nlines < 1 || t->ncols < 1 || t->line0 < 0 ||
        t->col0 < 0 || t->line0 + t->nlines > t->view.real_height ||
        t->col0 + t->ncols > t->view.real_width ||
        (uint64_t)t->nlines * (uint64_t)t->ncols > 0x7fffffffull)
        return MDZG_ERR_ARG;
    if (pal ? (!rgb || aa < 1 || aa > 64 || !mdzg_palette_ok(pal) || t->nlines % aa ||
               t->ncols % aa)
            : !t->counts)
        return MDZG_ERR_ARG;
    sreq* r = (sreq*)calloc(1, sizeof *r);
    if (!r)
        return MDZG_ERR_NOMEM;
    r->tile = *t;
    r->npix = (uint32_t)t->nlines * (uint32_t)t->ncols;
    r->slot = -1;
    if (pal) {
        r->image = 1;
        r->want_counts = t->counts != NULL;
        r->pal = *pal;
        r->ir.user_w = t->ncols / aa;
        r->ir.user_h = t->nlines / aa;
        r->ir.aa = aa;
        r->ir.pal = &r->pal;
        r->ir.rgb = rgb;
        r->ir.counts = t->counts;
        if (!r->tile.counts)
            r->tile.counts = rgb;   /* image jobs never scatter into it */
    }
    int rc = MDZG_OK;
    if (s->cfg.mode == MDZG_SERVE_BATCHED) {
        rc = mdzg_view_parse(&t->view, &r->pv);
        if (rc != MDZG_OK) {
            free(r);
            return rc;
        }
        r->parsed = 1;
        r->bytes = req_bytes(r);
        if (r->bytes > s->arena_cap) {      /* could never be admitted */
            mdzg_view_clear(&r->pv);
            free(r);
            return MDZG_ERR_BUSY;
        }
    }
    pthread_mutex_lock(&s->lock);
    if (s->shutdown) {
        pthread_mutex_unlock(&s->lock);
        if (r->parsed) mdzg_view_clear(&r->pv);
        free(r);
        return MDZG_ERR_STATE;
    }
    if (s->nreq == s->capreq) {
        s->capreq = s->capreq ? 2 * s->capreq : 256;
        s->reqs = (sreq**)realloc(s->reqs, sizeof(sreq*) * s->capreq);
    }
    if (s->pend_tail == s->cappend) {
        s->cappend = s->cappend ? 2 * s->cappend : 256;
        s->pend = (size_t*)realloc(s->pend, sizeof(size_t) * s->cappend);
    }
    r->id = s->nreq + 1;
    r->t_submit = mdzg_now_ms();
    s->reqs[s->nreq] = r;
    s->pend[s->pend_tail++] = s->nreq;
    s->nreq++;
    s->st.submitted++;
    *id = r->id;
    pthread_cond_signal(&s->cv_work);
    pthread_mutex_unlock(&s->lock);
    return rc;
}

int mdzg_server_submit(mdzg_server* s, const mdzg_tile* t, uint64_t* id)
{
    return submit(s, t, 0, NULL, NULL, id);
}

int mdzg_server_submit_image(mdzg_server* s, const mdzg_tile* t, int aa, const mdzg_palette* pal,
                             uint32_t* rgb, uint64_t* id)
{
    if (!pal)
        return MDZG_ERR_ARG;
    return submit(s, t, aa, pal, rgb, id);
}

static sreq* find_req(mdzg_server* s, uint64_t id)
{
    return (id >= 1 && id <= s->nreq) ? s->reqs[id - 1] : NULL;
}

int mdzg_server_poll(mdzg_server* s, uint64_t id)
{
    if (!s)
        return MDZG_ERR_ARG;
    pthread_mutex_lock(&s->lock);
    sreq* r = find_req(s, id);
    int st = !r ? MDZG_ERR_ARG : (r->status == 2 ? 1 : (r->status < 0 ? r->status : 0));
    pthread_mutex_unlock(&s->lock);
    return st;
}

int mdzg_server_wait(mdzg_server* s, uint64_t id, mdzg_req_info* info)
{
    if (!s)
        return MDZG_ERR_ARG;
    pthread_mutex_lock(&s->lock);
    sreq* r = find_req(s, id);
    if (!r) {
        pthread_mutex_unlock(&s->lock);
        return MDZG_ERR_ARG;
    }
    while (r->status == 0 || r->status == 1)
        pthread_cond_wait(&s->cv_done, &s->lock);
    if (info)
        fill_info(r, info);
    int rc = r->status == 2 ? MDZG_OK : r->status;
    pthread_mutex_unlock(&s->lock);
    return rc;
}

int mdzg_server_info(mdzg_server* s, uint64_t id, mdzg_req_info* info)
{
    if (!s || !info)
        return MDZG_ERR_ARG;
    pthread_mutex_lock(&s->lock);
    sreq* r = find_req(s, id);
    if (r)
        fill_info(r, info);
    pthread_mutex_unlock(&s->lock);
    return r ? MDZG_OK : MDZG_ERR_ARG;
}

int mdzg_server_stats(mdzg_server* s, mdzg_serve_stats* out)
{
    if (!s || !out)
        return MDZG_ERR_ARG;
    pthread_mutex_lock(&s->lock);
    *out = s->st;
    out->hbm_in_use = s->arena_used;
    out->hbm_peak = s->arena_peak;
    out->active = (uint64_t)s->nact;
    out->pending = (uint64_t)(s->pend_tail - s->pend_head);
    pthread_mutex_unlock(&s->lock);
    return MDZG_OK;
}

void mdzg_server_destroy(mdzg_server* s)
{
    if (!s)
        return;
    if (s->thr_started) {
        pthread_mutex_lock(&s->lock);
        s->shutdown = 1;
        pthread_cond_broadcast(&s->cv_work);
        pthread_mutex_unlock(&s->lock);
        pthread_join(s->thr, NULL);
    }
    hipSetDevice(s->ctx->device);
    for (size_t i = 0; i < s->nreq; ++i) {
        if (s->reqs[i]->parsed)
            mdzg_view_clear(&s->reqs[i]->pv);
        if (s->reqs[i]->job)
            mdzg_job_free(s->reqs[i]->job);
        free(s->reqs[i]);
    }
    free(s->reqs);
    free(s->pend);
    free(s->act);
    free(s->blocks);
    free(s->slot_used);
    for (int k = 0; k < MAX_KID; ++k) {
        if (s->streams[k]) hipStreamDestroy(s->streams[k]);
        if (s->ev_join[k]) hipEventDestroy(s->ev_join[k]);
    }
    if (s->ev_fork) hipEventDestroy(s->ev_fork);
    if (s->d_arena) hipFree(s->d_arena);
    if (s->d_remaining) hipFree(s->d_remaining);
    if (s->h_remaining) hipHostFree(s->h_remaining);
    if (s->d_ctr) hipFree(s->d_ctr);
    if (s->d_desc) hipFree(s->d_desc);
    if (s->h_desc) hipHostFree(s->h_desc);
    if (s->h_stage) hipHostFree(s->h_stage);
    pthread_mutex_destroy(&s->lock);
    pthread_cond_destroy(&s->cv_work);
    pthread_cond_destroy(&s->cv_done);
    free(s);
}
