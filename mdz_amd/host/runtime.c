/*
 * runtime.c -- device context, staging and launch of tile batches (C99).
 *
 * A batch (one mdzg_job) lays out in one device allocation:
 *   [descriptors][queue counters + cancel flag][start values][counts]
 * staged through one pinned host buffer: a single host-to-device copy before
 * the launches and a single device-to-host copy of the counts after them.
 * Tiles are grouped by arithmetic class; each class gets one persistent
 * launch draining its own queue (MDZG_BATCH_PERSISTENT), or each tile its own
 * plain launch (MDZG_BATCH_PER_TILE, the baseline form).  Jobs run on a pool
 * of streams so independent jobs overlap on the device.
 *
 * This replaces the pthread line pool of src/render_threads.c:260-393 for the
 * device: the claim of the next piece of work (next_line += aa under a mutex,
 * :366-369) becomes an atomic on a device counter, one pixel per lane.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "mdzg_host.h"

#define NSTREAMS 8
#define NCACHE 8
#define ALIGN_UP(x, a) (((x) + (size_t)(a) - 1) & ~((size_t)(a) - 1))

typedef struct cache_slot {
    void* d;
    size_t dcap;
    void* h;
    size_t hcap;
    int used;
} cache_slot;

struct mdzg_rt {
    hipStream_t streams[NSTREAMS];
    int next_stream;
    cache_slot cache[NCACHE];
    pthread_mutex_t cache_lock;
};

/* the ctx struct of mdzg_host.h plus the runtime pools */
typedef struct ctx_full {
    struct mdzg_ctx c;
    struct mdzg_rt rt;
} ctx_full;

struct mdzg_job {
    mdzg_ctx* ctx;
    int ntiles;
    mdzg_tile* tiles;            /* copies of the caller's tile descriptors */
    int mode;
    hipStream_t stream;
    int slot;                    /* cache slot index, or -1 when owned */
    void* d;
    void* h;
    size_t off_ctr, off_coords, off_out, bytes, out_bytes;
    size_t* out_off;
    int nkid;
    int kids[MDZG_NKERNELS];
    hipEvent_t ev0, ev1;
    hipEvent_t ev_fork;
    hipEvent_t ev_cls[MDZG_NKERNELS];
    double t_start;
    int launched;
    int done;
    int rc;
    int launches;
    int cancelled;
    /* image jobs (mdzg_render_image): colour the frame on the device */
    const mdzg_image_req* img;
    size_t off_pal, off_iters, off_rgb, rgb_bytes;
};

double mdzg_now_ms(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return 1e3 * (double)ts.tv_sec + 1e-6 * (double)ts.tv_nsec;
}

int mdzg_abi_version(void) { return MDZG_ABI_VERSION; }

const char* mdzg_strerror(int code)
{
    switch (code) {
    case MDZG_OK: return "ok";
    case MDZG_ERR_ARG: return "invalid argument";
    case MDZG_ERR_NOMEM: return "out of memory";
    case MDZG_ERR_DEVICE: return "HIP device error";
    case MDZG_ERR_STOPPED: return "render stopped";
    case MDZG_ERR_BUSY: return "admission refused";
    case MDZG_ERR_PARSE: return "coordinate string did not parse";
    case MDZG_ERR_RANGE: return "precision or size out of range";
    case MDZG_ERR_STATE: return "call out of sequence";
    default: return "unknown error";
    }
}

/* iterations a lane runs between queue refills, by kernel class (measured on
   the whole set, centre -0.75, width 3, at 240x180 and 3840x2160): the FP64
   and 32/64-bit RN loops are cheap enough per iteration that a refill round
   every iteration costs more than the lanes a longer chunk leaves idle.  Wider
   words run 8 (a 640x480 frame at 128 bits +3.7 % against 1, a 1024-bit band
   +1.6 % at 4): a few idle iterations per finished pixel cost less than a
   round per iteration; the GMP loops too (+1.3 % at 128 bits).  The FP64 loop
   leaves a run early once all its lanes have escaped; 256 iterations between
   rounds took the 3840x2160 frame 1.35 ms against 1.41 ms at 128 and 512,
   with the 240x180 frame and an all-interior frame within 2 % of their
   best. */
static int default_chunk(int kid)
{
    if (kid == 0)
        return 256;
    if (kid >= 1 && kid < MDZG_KID_GMP0)
        return mdzg_nwords[(kid - 1) >> 1] <= 2 ? 16 : 8;
    return 8;
}

int mdzg_create(int device, mdzg_ctx** out)
{
    if (!out)
        return MDZG_ERR_ARG;
    *out = NULL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return MDZG_ERR_DEVICE;
    if (hipSetDevice(device) != hipSuccess)
        return MDZG_ERR_DEVICE;
    ctx_full* f = (ctx_full*)calloc(1, sizeof *f);
    if (!f)
        return MDZG_ERR_NOMEM;
    mdzg_ctx* c = &f->c;
    c->device = device;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess)
        goto fail;
    c->ncu = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->ctl, hipStreamNonBlocking) != hipSuccess)
        goto fail;
    for (int i = 0; i < NSTREAMS; ++i)
        if (hipStreamCreateWithFlags(&f->rt.streams[i], hipStreamNonBlocking) != hipSuccess)
            goto fail;
    if (hipMalloc((void**)&c->d_cancel_zero, 256) != hipSuccess)
        goto fail;
    if (hipMemset(c->d_cancel_zero, 0, 256) != hipSuccess)
        goto fail;
    /* the no-allocation entry point's scratch: its device address on this
       device, a host staging buffer of the same size, two timing events */
    if (mdzg_dev_scratch(device, &c->d_scratch, NULL) != 0)
        goto fail;
    c->h_scratch = malloc(MDZG_NOALLOC_SCRATCH_BYTES);
    if (!c->h_scratch || hipEventCreate(&c->na_ev0) != hipSuccess ||
        hipEventCreate(&c->na_ev1) != hipSuccess)
        goto fail;
    for (int k = 0; k < mdzg_dev_nkernels(); ++k) {
        c->occ[k] = mdzg_dev_occupancy(k, 1);
        if (c->occ[k] < 1)
            c->occ[k] = 1;
        c->chunk[k] = default_chunk(k);
    }
    c->graphs = 1;
    {
        const char* fp = getenv("MDZG_FAILPOINT_PLAN");
        c->fail_plan = (fp && *fp) ? strtol(fp, NULL, 10) : 0;
    }
    c->band_pixels = MDZG_DEFAULT_BAND_PIXELS;
    c->band_inflight = MDZG_DEFAULT_BAND_INFLIGHT;
    c->gc = mdzg_gcache_create();
    if (!c->gc)
        goto fail;
    pthread_mutex_init(&c->lock, NULL);
    pthread_mutex_init(&f->rt.cache_lock, NULL);
    /* the tuned table: MDZG_TUNE_TABLE, else tune_gfx950.txt beside the library */
    {
        const char* env = getenv("MDZG_TUNE_TABLE");
        char path[4096];
        if (env && *env) {
            snprintf(path, sizeof path, "%s", env);
            mdzg_tune_load(c, path);
        } else if (mdzg_lib_dir(path, sizeof path) == 0) {
            size_t n = strlen(path);
            snprintf(path + n, sizeof path - n, "/tune_gfx950.txt");
            mdzg_tune_load(c, path);
        }
    }
    *out = c;
    return MDZG_OK;
fail:
    free(c->h_scratch);
    free(f);
    return MDZG_ERR_DEVICE;
}

void mdzg_destroy(mdzg_ctx* c)
{
    if (!c)
        return;
    ctx_full* f = (ctx_full*)c;
    hipSetDevice(c->device);
    hipDeviceSynchronize();
    for (int i = 0; i < NCACHE; ++i) {
        if (f->rt.cache[i].d) hipFree(f->rt.cache[i].d);
        if (f->rt.cache[i].h) hipHostFree(f->rt.cache[i].h);
    }
    for (int i = 0; i < NSTREAMS; ++i)
        if (f->rt.streams[i]) hipStreamDestroy(f->rt.streams[i]);
    if (c->stream) hipStreamDestroy(c->stream);
    if (c->ctl) hipStreamDestroy(c->ctl);
    if (c->d_cancel_zero) hipFree(c->d_cancel_zero);
    mdzg_gcache_destroy(c->gc);
    if (c->na_ev0) hipEventDestroy(c->na_ev0);
    if (c->na_ev1) hipEventDestroy(c->na_ev1);
    free(c->h_scratch);
    pthread_mutex_destroy(&c->lock);
    pthread_mutex_destroy(&f->rt.cache_lock);
    free(f);
}

/* ---- buffer cache -------------------------------------------------------- */
static int acquire(mdzg_ctx* c, size_t bytes, void** d, void** h)
{
    ctx_full* f = (ctx_full*)c;
    pthread_mutex_lock(&f->rt.cache_lock);
    int best = -1;
    for (int i = 0; i < NCACHE; ++i) {
        cache_slot* s = &f->rt.cache[i];
        if (!s->used && s->d && s->dcap >= bytes && (best < 0 || s->dcap < f->rt.cache[best].dcap))
            best = i;
    }
    if (best < 0) {
        /* grow a free slot (or take an empty one) */
        for (int i = 0; i < NCACHE; ++i) {
            cache_slot* s = &f->rt.cache[i];
            if (!s->used) {
                if (s->d) { hipFree(s->d); s->d = NULL; }
                if (s->h) { hipHostFree(s->h); s->h = NULL; }
                size_t cap = ALIGN_UP(bytes + bytes / 4, 1 << 20);
                if (hipMalloc(&s->d, cap) != hipSuccess) { s->d = NULL; break; }
                if (hipHostMalloc(&s->h, cap, hipHostMallocDefault) != hipSuccess) {
                    hipFree(s->d);
                    s->d = NULL;
                    s->h = NULL;
                    break;
                }
                s->dcap = s->hcap = cap;
                best = i;
                break;
            }
        }
    }
    if (best >= 0) {
        f->rt.cache[best].used = 1;
        *d = f->rt.cache[best].d;
        *h = f->rt.cache[best].h;
    }
    pthread_mutex_unlock(&f->rt.cache_lock);
    return best;
}

static void release(mdzg_ctx* c, int slot)
{
    ctx_full* f = (ctx_full*)c;
    pthread_mutex_lock(&f->rt.cache_lock);
    if (slot >= 0 && slot < NCACHE)
        f->rt.cache[slot].used = 0;
    pthread_mutex_unlock(&f->rt.cache_lock);
}

static hipStream_t pick_stream(mdzg_ctx* c)
{
    ctx_full* f = (ctx_full*)c;
    pthread_mutex_lock(&f->rt.cache_lock);
    hipStream_t s = f->rt.streams[f->rt.next_stream];
    f->rt.next_stream = (f->rt.next_stream + 1) % NSTREAMS;
    pthread_mutex_unlock(&f->rt.cache_lock);
    return s;
}

/* the k-th stream of the pool after `main` (distinct from it for k < NSTREAMS) */
static hipStream_t side_stream(mdzg_ctx* c, hipStream_t main, int k)
{
    ctx_full* f = (ctx_full*)c;
    int base = 0;
    for (int i = 0; i < NSTREAMS; ++i)
        if (f->rt.streams[i] == main)
            base = i;
    return f->rt.streams[(base + k) % NSTREAMS];
}

/* ---- tuning --------------------------------------------------------------- */
int mdzg_tune_set(mdzg_ctx* c, int kernel_id, int blocks_per_cu, int chunk)
{
    if (!c || kernel_id < 0 || kernel_id >= mdzg_dev_nkernels() || blocks_per_cu < 0 ||
        chunk < 0 || chunk > 4096)
        return MDZG_ERR_ARG;
    if (blocks_per_cu > 0) {
        int cap = mdzg_dev_occupancy(kernel_id, 1);
        c->occ[kernel_id] = blocks_per_cu < cap ? blocks_per_cu : (cap > 0 ? cap : 1);
    }
    if (chunk > 0)
        c->chunk[kernel_id] = chunk;
    return MDZG_OK;
}

int mdzg_tune_get(mdzg_ctx* c, int kernel_id, int* blocks_per_cu, int* chunk, int* max_blocks_per_cu)
{
    if (!c || kernel_id < 0 || kernel_id >= mdzg_dev_nkernels())
        return MDZG_ERR_ARG;
    if (blocks_per_cu) *blocks_per_cu = c->occ[kernel_id];
    if (chunk) *chunk = c->chunk[kernel_id];
    if (max_blocks_per_cu) *max_blocks_per_cu = mdzg_dev_occupancy(kernel_id, 1);
    return MDZG_OK;
}

int mdzg_kernel_id(int path, long precision)
{
    int nw, kclr;
    return mdzg_class_of(path, precision, &nw, &kclr);
}

/* ---- job construction ---------------------------------------------------- */
static int tile_ok(const mdzg_tile* t)
{
    return t && t->counts && t->nlines > 0 && t->ncols > 0 && t->line0 >= 0 && t->col0 >= 0 &&
           t->line0 + t->nlines <= t->view.real_height && t->col0 + t->ncols <= t->view.real_width &&
           (uint64_t)t->nlines * (uint64_t)t->ncols <= 0x7fffffffull;
}

static void job_free_mem(mdzg_job* j)
{
    if (j->ev0) hipEventDestroy(j->ev0);
    if (j->ev1) hipEventDestroy(j->ev1);
    if (j->ev_fork) hipEventDestroy(j->ev_fork);
    for (int k = 0; k < MDZG_NKERNELS; ++k)
        if (j->ev_cls[k]) hipEventDestroy(j->ev_cls[k]);
    if (j->slot >= 0) {
        release(j->ctx, j->slot);
    } else {
        if (j->d) hipFree(j->d);
        if (j->h) hipHostFree(j->h);
    }
    free(j->tiles);
    free(j->out_off);
    free(j);
}

static int mdzg_job_start(mdzg_ctx* c, const mdzg_tile* tiles, int ntiles, int mode,
                          const mdzg_image_req* img, int sync, mdzg_job** out)
{
    if (!c || !tiles || ntiles < 1 || !out || (mode != MDZG_BATCH_PERSISTENT && mode != MDZG_BATCH_PER_TILE))
        return MDZG_ERR_ARG;
    *out = NULL;
    for (int i = 0; i < ntiles; ++i)
        if (!tile_ok(&tiles[i]))
            return MDZG_ERR_ARG;

    mdzg_job* j = (mdzg_job*)calloc(1, sizeof *j);
    if (!j)
        return MDZG_ERR_NOMEM;
    j->ctx = c;
    j->slot = -1;
    j->mode = mode;
    j->img = img;
    j->ntiles = ntiles;
    j->tiles = (mdzg_tile*)malloc(sizeof(mdzg_tile) * (size_t)ntiles);
    j->out_off = (size_t*)malloc(sizeof(size_t) * (size_t)ntiles);
    mdzg_pview* pv = (mdzg_pview*)calloc((size_t)ntiles, sizeof(mdzg_pview));
    int* order = (int*)malloc(sizeof(int) * (size_t)ntiles);
    if (!j->tiles || !j->out_off || !pv || !order) {
        free(pv);
        free(order);
        job_free_mem(j);
        return MDZG_ERR_NOMEM;
    }
    memcpy(j->tiles, tiles, sizeof(mdzg_tile) * (size_t)ntiles);
    hipSetDevice(c->device);

    int rc = MDZG_OK;
    int nparsed = 0;
    for (int i = 0; i < ntiles && rc == MDZG_OK; ++i, ++nparsed)
        rc = mdzg_view_parse(&tiles[i].view, &pv[i]);
    if (rc != MDZG_OK) {
        nparsed--;   /* the failing parse cleaned up after itself */
        goto done;
    }

    /* order tiles by kernel id (stable) */
    {
        int n = 0;
        for (int k = 0; k < mdzg_dev_nkernels(); ++k) {
            int any = 0;
            for (int i = 0; i < ntiles; ++i)
                if (pv[i].kid == k) {
                    order[n++] = i;
                    any = 1;
                }
            if (any)
                j->kids[j->nkid++] = k;
        }
    }

    /* layout */
    size_t off = 0;
    const size_t off_desc = 0;
    off = ALIGN_UP(sizeof(mdzg_dtile) * (size_t)ntiles, 256);
    j->off_ctr = off;
    off = ALIGN_UP(off + 8 * (size_t)(j->nkid + 1), 256);   /* counters + cancel flag */
    j->off_iters = off;                                      /* image: pixel-iteration sum */
    j->off_pal = off + 256;                                  /* image: palette */
    if (img)
        off += 256 + 4 * 256;
    j->off_coords = off;
    for (int i = 0; i < ntiles; ++i) {
        if (pv[i].vs) {
            size_t nval = (size_t)tiles[i].ncols + (size_t)tiles[i].nlines + 2;
            off += ALIGN_UP(nval * 4 * (size_t)pv[i].vs, 16);
        }
    }
    off = ALIGN_UP(off, 256);
    j->off_out = off;
    for (int i = 0; i < ntiles; ++i) {
        j->out_off[i] = off;
        off += ALIGN_UP((size_t)tiles[i].nlines * (size_t)tiles[i].ncols * 4, 16);
    }
    off = ALIGN_UP(off, 256);
    j->out_bytes = off - j->off_out;
    j->off_rgb = off;
    if (img) {
        j->rgb_bytes = 4 * (size_t)img->user_w * (size_t)img->user_h;
        off += j->rgb_bytes;
    }
    j->bytes = ALIGN_UP(off, 256);

    j->slot = acquire(c, j->bytes, &j->d, &j->h);
    if (j->slot < 0) {
        /* every cached buffer is in use: this job owns its memory */
        if (hipMalloc(&j->d, j->bytes) != hipSuccess) {
            j->d = NULL;
            rc = MDZG_ERR_NOMEM;
            goto done;
        }
        if (hipHostMalloc(&j->h, j->bytes, hipHostMallocDefault) != hipSuccess) {
            j->h = NULL;
            rc = MDZG_ERR_NOMEM;
            goto done;
        }
    }

    /* fill staging: descriptors, counters, start values */
    {
        char* H = (char*)j->h;
        char* D = (char*)j->d;
        memset(H, 0, j->off_out);
        if (img)
            memcpy(H + j->off_pal, img->pal->colours, 4 * (size_t)img->pal->n);
        mdzg_dtile* desc = (mdzg_dtile*)(H + off_desc);
        size_t coff = j->off_coords;
        uint64_t pix_begin = 0;
        int prev_kid = -1;
        for (int n = 0; n < ntiles; ++n) {
            const int i = order[n];
            const mdzg_tile* t = &tiles[i];
            mdzg_dtile* d = &desc[n];
            if (pv[i].kid != prev_kid) {
                pix_begin = 0;
                prev_kid = pv[i].kid;
            }
            d->out = (uint32_t*)(D + j->out_off[i]);
            d->ncols = t->ncols;
            d->nrows = t->nlines;
            d->col0 = t->col0;
            d->line0 = t->line0;
            d->W = t->view.real_width;
            d->fractal = t->view.fractal;
            d->family = t->view.family;
            d->kclr = pv[i].kclr;
            d->depth = (int32_t)t->view.depth;
            d->vs = pv[i].vs;
            d->gprec = pv[i].v.path == MDZG_PATH_GMP ? pv[i].nw : 0;
            d->pix_begin = pix_begin;
            pix_begin += (uint64_t)t->ncols * (uint64_t)t->nlines;
            if (pv[i].vs) {
                const size_t vsb = 4 * (size_t)pv[i].vs;
                uint32_t* xs = (uint32_t*)(H + coff);
                uint32_t* ys = (uint32_t*)(H + coff + vsb * (size_t)t->ncols);
                uint32_t* jc = (uint32_t*)(H + coff + vsb * ((size_t)t->ncols + (size_t)t->nlines));
                rc = mdzg_coords_fill(&pv[i], t->line0, t->nlines, t->col0, t->ncols, xs, ys, jc);
                if (rc != MDZG_OK)
                    goto done;
                d->xs = (const uint32_t*)(D + coff);
                d->ys = (const uint32_t*)(D + coff + vsb * (size_t)t->ncols);
                d->jc = (const uint32_t*)(D + coff + vsb * ((size_t)t->ncols + (size_t)t->nlines));
                coff += ALIGN_UP(((size_t)t->ncols + (size_t)t->nlines + 2) * vsb, 16);
            } else {
                mdzg_coords_fp64(&pv[i], &d->xmin, &d->ymax, &d->width, &d->jre, &d->jim);
            }
        }
    }

    /* copy in, launch, copy out */
    {
        char* D = (char*)j->d;
        const mdzg_dtile* d_desc = (const mdzg_dtile*)(D + off_desc);
        unsigned long long* d_ctr = (unsigned long long*)(D + j->off_ctr);
        const int* d_cancel = (const int*)(D + j->off_ctr + 8 * (size_t)j->nkid);
        j->stream = pick_stream(c);
        if (hipEventCreate(&j->ev0) != hipSuccess || hipEventCreate(&j->ev1) != hipSuccess) {
            rc = MDZG_ERR_DEVICE;
            goto done;
        }
        if (hipMemcpyAsync(j->d, j->h, j->off_out, hipMemcpyHostToDevice, j->stream) != hipSuccess) {
            rc = MDZG_ERR_DEVICE;
            goto done;
        }
        /* the counts of an asynchronous job, which can be cancelled, start at
           the UINT_MAX sentinel of src/image_info.c:214-215, so that the pixels
           it never finished stay recognisable: a fill kernel, the root of the
           launch plan.  A synchronous render cannot be cancelled and its
           kernels write every count, so it needs no fill. */
        mdzg_gnode fillnode, tail;
        mdzg_gnode* fill = sync ? NULL : &fillnode;
        mdzg_gnode par[MDZG_NKERNELS];
        int npar = 0;
        if (fill && mdzg_dev_node_fill(fill, (uint32_t*)(D + j->off_out), j->out_bytes / 4,
                                       0xffffffffu) != 0) {
            rc = MDZG_ERR_DEVICE;
            goto done;
        }
        int first = 0;
        for (int k = 0; k < j->nkid; ++k) {
            const int kid = j->kids[k];
            int cnt = 0;
            uint64_t total = 0;
            while (first + cnt < ntiles && pv[order[first + cnt]].kid == kid) {
                const mdzg_tile* t = &tiles[order[first + cnt]];
                total += (uint64_t)t->ncols * (uint64_t)t->nlines;
                ++cnt;
            }
            if (mode == MDZG_BATCH_PERSISTENT) {
                /* each arithmetic class drains its own queue; the class
                   launches run concurrently after the fill */
                uint64_t cap = (uint64_t)c->ncu * (uint64_t)c->occ[kid];
                uint64_t need = (total + 255) / 256;
                int grid = (int)(need < cap ? need : cap);
                if (mdzg_dev_node_persist(&par[npar++], kid, d_desc + first, cnt, total, d_ctr + k,
                                          d_cancel, grid, c->chunk[kid]) != 0) {
                    rc = MDZG_ERR_DEVICE;
                    goto done;
                }
            }
            first += cnt;
        }
        const int has_tail = img != NULL;
        if (img) {
            /* one tile, the whole frame: colour it where it lies */
            const mdzg_palette* P = img->pal;
            if (mdzg_dev_node_colour(&tail, (const uint32_t*)(D + j->out_off[0]), tiles[0].ncols,
                                     img->user_w, img->user_h, img->aa,
                                     (const uint32_t*)(D + j->off_pal), P->n, P->offset,
                                     P->colour_scale, P->interpolate,
                                     (uint32_t)tiles[0].view.depth, (uint32_t*)(D + j->off_rgb),
                                     (unsigned long long*)(D + j->off_iters),
                                     MDZG_COLOUR_BLOCKS_PER_CU * c->ncu) != 0) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
        }
        hipEventRecord(j->ev0, j->stream);
        if (c->fail_plan > 0 && __atomic_add_fetch(&c->plans, 1, __ATOMIC_RELAXED) == c->fail_plan) {
            rc = MDZG_ERR_DEVICE;
            goto done;
        }
        if (mode == MDZG_BATCH_PERSISTENT && c->graphs && npar >= 2) {
            /* concurrent class launches: the whole plan as one replay of a
               cached graph (a chain of launches on one stream goes out
               directly: on this stack a graph launch of one kernel took the
               240x180 FP64 frame from 40 to 48 us) */
            if (mdzg_graph_run(c->gc, fill, par, npar, has_tail ? &tail : NULL, j->stream) != 0) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
            j->launches += (fill != NULL) + npar + has_tail;
        } else if (mode == MDZG_BATCH_PERSISTENT) {
            /* launched one by one: the class launches fork onto separate
               streams so their tails overlap, and join back */
            if (fill && mdzg_dev_node_launch(fill, j->stream) != 0) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
            const int fork = npar > 1;
            if (fork) {
                if (hipEventCreateWithFlags(&j->ev_fork, hipEventDisableTiming) != hipSuccess) {
                    rc = MDZG_ERR_DEVICE;
                    goto done;
                }
                hipEventRecord(j->ev_fork, j->stream);
            }
            for (int k = 0; k < npar; ++k) {
                hipStream_t st = j->stream;
                if (fork && k > 0) {
                    st = side_stream(c, j->stream, k);
                    hipStreamWaitEvent(st, j->ev_fork, 0);
                }
                if (mdzg_dev_node_launch(&par[k], st) != 0) {
                    rc = MDZG_ERR_DEVICE;
                    goto done;
                }
                if (fork && k > 0) {
                    if (hipEventCreateWithFlags(&j->ev_cls[k], hipEventDisableTiming) !=
                        hipSuccess) {
                        rc = MDZG_ERR_DEVICE;
                        goto done;
                    }
                    hipEventRecord(j->ev_cls[k], st);
                }
            }
            if (fork)
                for (int k = 1; k < npar; ++k)
                    hipStreamWaitEvent(j->stream, j->ev_cls[k], 0);
            if (has_tail && mdzg_dev_node_launch(&tail, j->stream) != 0) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
            j->launches += (fill != NULL) + npar + has_tail;
        } else {
            /* the per-tile baseline: one plain launch per tile, in order */
            if (fill) {
                if (mdzg_dev_node_launch(fill, j->stream) != 0) {
                    rc = MDZG_ERR_DEVICE;
                    goto done;
                }
                j->launches++;
            }
            int f0 = 0;
            for (int k = 0; k < j->nkid; ++k) {
                const int kid = j->kids[k];
                int cnt = 0;
                while (f0 + cnt < ntiles && pv[order[f0 + cnt]].kid == kid)
                    ++cnt;
                for (int n = 0; n < cnt; ++n) {
                    const mdzg_tile* t = &tiles[order[f0 + n]];
                    uint32_t npix = (uint32_t)t->ncols * (uint32_t)t->nlines;
                    if (mdzg_dev_launch_tile(kid, d_desc + f0 + n, npix, j->stream) != 0) {
                        rc = MDZG_ERR_DEVICE;
                        goto done;
                    }
                    j->launches++;
                }
                f0 += cnt;
            }
            if (has_tail && mdzg_dev_node_launch(&tail, j->stream) != 0) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
            j->launches += has_tail;
        }
        hipEventRecord(j->ev1, j->stream);
        if (img) {
            if (hipMemcpyAsync((char*)j->h + j->off_iters, D + j->off_iters, 8,
                               hipMemcpyDeviceToHost, j->stream) != hipSuccess ||
                hipMemcpyAsync((char*)j->h + j->off_rgb, D + j->off_rgb, j->rgb_bytes,
                               hipMemcpyDeviceToHost, j->stream) != hipSuccess ||
                (img->counts && hipMemcpyAsync((char*)j->h + j->off_out, D + j->off_out,
                                               j->out_bytes, hipMemcpyDeviceToHost,
                                               j->stream) != hipSuccess)) {
                rc = MDZG_ERR_DEVICE;
                goto done;
            }
        } else if (hipMemcpyAsync((char*)j->h + j->off_out, D + j->off_out, j->out_bytes,
                                  hipMemcpyDeviceToHost, j->stream) != hipSuccess) {
            rc = MDZG_ERR_DEVICE;
            goto done;
        }
        j->launched = 1;
    }

done:
    for (int i = 0; i < nparsed; ++i)
        mdzg_view_clear(&pv[i]);
    free(pv);
    free(order);
    if (rc != MDZG_OK) {
        /* nothing the job launched may still use its memory: wait for its
           stream and for the side streams its class launches forked onto */
        if (j->launched || j->stream) {
            ctx_full* f = (ctx_full*)c;
            for (int i = 0; i < NSTREAMS; ++i)
                hipStreamSynchronize(f->rt.streams[i]);
        }
        job_free_mem(j);
        return rc;
    }
    *out = j;
    return MDZG_OK;
}

int mdzg_submit(mdzg_ctx* c, const mdzg_tile* tiles, int ntiles, mdzg_job** out)
{
    int rc = mdzg_job_start(c, tiles, ntiles, MDZG_BATCH_PERSISTENT, NULL, 0, out);
    if (rc == MDZG_OK)
        (*out)->t_start = mdzg_now_ms();
    return rc;
}

int mdzg_rt_submit_image(mdzg_ctx* c, const mdzg_tile* tile, const mdzg_image_req* img,
                         mdzg_job** out)
{
    int rc = mdzg_job_start(c, tile, 1, MDZG_BATCH_PERSISTENT, img, 0, out);
    if (rc == MDZG_OK)
        (*out)->t_start = mdzg_now_ms();
    return rc;
}

int mdzg_poll(mdzg_job* j)
{
    if (!j)
        return MDZG_ERR_ARG;
    if (j->done)
        return 1;
    hipError_t e = hipStreamQuery(j->stream);
    if (e == hipSuccess)
        return 1;
    if (e == hipErrorNotReady)
        return 0;
    return MDZG_ERR_DEVICE;
}

int mdzg_cancel(mdzg_job* j)
{
    if (!j || !j->launched)
        return MDZG_ERR_ARG;
    mdzg_ctx* c = j->ctx;
    int* d_cancel = (int*)((char*)j->d + j->off_ctr + 8 * (size_t)j->nkid);
    hipSetDevice(c->device);
    static const int one = 1;
    if (hipMemcpyAsync(d_cancel, &one, sizeof one, hipMemcpyHostToDevice, c->ctl) != hipSuccess)
        return MDZG_ERR_DEVICE;
    hipStreamSynchronize(c->ctl);
    j->cancelled = 1;
    return MDZG_OK;
}

/* dst = src (n counts) and their pixel-iterations, a zero count standing for
   the depth: one pass with integer sums (exact, like the double sum of each
   term, below 2^53), which the compiler vectorises.  A double accumulator
   carried one 4-cycle add per pixel: 0.3 ms of host time per 640x480 frame. */
__attribute__((optimize("O3"))) double mdzg_take_counts(uint32_t* dst, const uint32_t* src,
                                                        size_t n, double depth)
{
    uint64_t sum = 0, zeros = 0;
    for (size_t k = 0; k < n; ++k) {
        const uint32_t v = src[k];
        dst[k] = v;
        sum += v;
        zeros += (v == 0);
    }
    return (double)sum + (double)zeros * depth;
}

int mdzg_wait(mdzg_job* j, mdzg_stats* st)
{
    if (!j)
        return MDZG_ERR_ARG;
    if (!j->done) {
        if (hipStreamSynchronize(j->stream) != hipSuccess)
            return MDZG_ERR_DEVICE;
        /* scatter counts to the callers' buffers */
        double pi = 0.0;
        if (j->img) {
            memcpy(j->img->rgb, (char*)j->h + j->off_rgb, j->rgb_bytes);
            if (j->img->counts)
                memcpy(j->img->counts, (char*)j->h + j->off_out,
                       4 * (size_t)j->tiles[0].nlines * (size_t)j->tiles[0].ncols);
            pi = (double)*(const unsigned long long*)((char*)j->h + j->off_iters);
        }
        for (int i = 0; i < j->ntiles && !j->img; ++i) {
            const mdzg_tile* t = &j->tiles[i];
            const uint32_t* src = (const uint32_t*)((char*)j->h + j->out_off[i]);
            size_t n = (size_t)t->nlines * (size_t)t->ncols;
            const double depth = (double)t->view.depth;
            if (!j->cancelled) {
                pi += mdzg_take_counts(t->counts, src, n, depth);
            } else {
                /* keep the caller's contents where the device never wrote */
                for (size_t k = 0; k < n; ++k)
                    if (src[k] != 0xffffffffu) {
                        t->counts[k] = src[k];
                        pi += src[k] ? (double)src[k] : depth;
                    }
            }
        }
        float ms = 0.f;
        hipEventElapsedTime(&ms, j->ev0, j->ev1);
        j->done = 1;
        j->rc = j->cancelled ? MDZG_ERR_STOPPED : MDZG_OK;
        if (st) {
            st->kernel_ms = ms;
            st->wall_ms = mdzg_now_ms() - j->t_start;
            st->pixel_iters = pi;
            st->launches = j->launches;
        }
    }
    return j->rc;
}

void mdzg_job_free(mdzg_job* j)
{
    if (!j)
        return;
    if (j->launched && !j->done)
        hipStreamSynchronize(j->stream);
    job_free_mem(j);
}

int mdzg_render_batch(mdzg_ctx* c, const mdzg_tile* tiles, int ntiles, int mode, mdzg_stats* st)
{
    double t0 = mdzg_now_ms();
    mdzg_job* j = NULL;
    int rc = mdzg_job_start(c, tiles, ntiles, mode, NULL, 1, &j);
    if (rc != MDZG_OK)
        return rc;
    j->t_start = t0;
    rc = mdzg_wait(j, st);
    mdzg_job_free(j);
    return rc;
}

int mdzg_palette_ok(const mdzg_palette* p)
{
    return p && p->n >= 2 && p->n <= 256 && p->offset >= 0 && p->colour_scale > 0.0 &&
           p->colour_scale <= 1e300;
}

int mdzg_colour(mdzg_ctx* c, const uint32_t* counts, int user_w, int user_h, int aa,
                const mdzg_palette* pal, uint32_t* rgb, mdzg_stats* stats)
{
    const double t0 = mdzg_now_ms();
    if (!c || !counts || !rgb || user_w < 1 || user_h < 1 || aa < 1 || aa > 64 ||
        !mdzg_palette_ok(pal))
        return MDZG_ERR_ARG;
    const uint64_t ns = (uint64_t)user_w * aa * (uint64_t)user_h * aa;
    const uint64_t np = (uint64_t)user_w * (uint64_t)user_h;
    if (ns > 0x7fffffffull || (uint64_t)user_w * aa > 0x7fffffffull)
        return MDZG_ERR_RANGE;
    const size_t cb = ALIGN_UP(4 * (size_t)ns, 256), pb = 1024, ob = 4 * (size_t)np;
    const size_t bytes = cb + pb + ob;
    void *d = NULL, *h = NULL;
    hipSetDevice(c->device);
    int slot = acquire(c, bytes, &d, &h);
    if (slot < 0) {
        if (hipMalloc(&d, bytes) != hipSuccess)
            return MDZG_ERR_NOMEM;
        if (hipHostMalloc(&h, bytes, hipHostMallocDefault) != hipSuccess) {
            hipFree(d);
            return MDZG_ERR_NOMEM;
        }
    }
    memcpy(h, counts, 4 * (size_t)ns);
    memcpy((char*)h + cb, pal->colours, 4 * (size_t)pal->n);
    hipStream_t st = pick_stream(c);
    hipEvent_t e0 = NULL, e1 = NULL;
    int rc = MDZG_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
        hipMemcpyAsync(d, h, cb + pb, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipEventRecord(e0, st) != hipSuccess ||
        mdzg_dev_launch_colour((const uint32_t*)d, user_w * aa, user_w, user_h, aa,
                               (const uint32_t*)((char*)d + cb), pal->n, pal->offset,
                               pal->colour_scale, pal->interpolate, 0u,
                               (uint32_t*)((char*)d + cb + pb), NULL,
                               MDZG_COLOUR_BLOCKS_PER_CU * c->ncu, st) != 0 ||
        hipEventRecord(e1, st) != hipSuccess ||
        hipMemcpyAsync((char*)h + cb + pb, (char*)d + cb + pb, ob, hipMemcpyDeviceToHost,
                       st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        rc = MDZG_ERR_DEVICE;
    } else {
        memcpy(rgb, (char*)h + cb + pb, ob);
        if (stats) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, e0, e1);
            stats->kernel_ms = ms;
            stats->wall_ms = mdzg_now_ms() - t0;
            stats->pixel_iters = 0.0;
            stats->launches = 1;
        }
    }
    if (rc != MDZG_OK)
        hipStreamSynchronize(st);      /* nothing may still use the buffers */
    if (e0)
        hipEventDestroy(e0);
    if (e1)
        hipEventDestroy(e1);
    if (slot >= 0) {
        release(c, slot);
    } else {
        hipFree(d);
        hipHostFree(h);
    }
    return rc;
}

int mdzg_rt_render_image(mdzg_ctx* c, const mdzg_tile* tile, const mdzg_image_req* img,
                         mdzg_stats* st)
{
    double t0 = mdzg_now_ms();
    mdzg_job* j = NULL;
    int rc = mdzg_job_start(c, tile, 1, MDZG_BATCH_PERSISTENT, img, 1, &j);
    if (rc != MDZG_OK)
        return rc;
    j->t_start = t0;
    rc = mdzg_wait(j, st);
    mdzg_job_free(j);
    return rc;
}

int mdzg_render_tile(mdzg_ctx* c, const mdzg_tile* tile, mdzg_stats* st)
{
    return mdzg_render_batch(c, tile, 1, MDZG_BATCH_PERSISTENT, st);
}

/* ---- rendering without allocating ---------------------------------------
   mdzg_render_tile's work for callers that cannot allocate device memory:
   the counts go straight to the caller's device buffer, and the descriptor,
   queue counter, cancel flag and start values to the static scratch array
   of noalloc.hip (its device address c->d_scratch), staged through the
   context's host buffer in one copy.  Nothing is allocated: the events and
   the host buffer were made by mdzg_create.  One lock per device serialises
   the calls of this process, which share that device's array. */
#define NA_OFF_CTR 256                    /* queue counter (8 B), cancel flag (+8) */
#define NA_OFF_COORDS 512                 /* start values                          */
#define NA_MAX_DEVICES 64
static pthread_mutex_t na_lock[NA_MAX_DEVICES] = { [0 ... NA_MAX_DEVICES - 1] =
                                                       PTHREAD_MUTEX_INITIALIZER };

/* bytes of a one-tile staging area: descriptor, queue counter and cancel
   flag, then the start values */
static size_t stage_bytes(const mdzg_tile* t, const mdzg_pview* pv)
{
    const size_t vsb = 4 * (size_t)pv->vs;
    return NA_OFF_COORDS +
           (pv->vs ? ALIGN_UP(((size_t)t->ncols + (size_t)t->nlines + 2) * vsb, 16) : 0);
}

/* the tile's descriptor and start values into the host image H of a
   staging area whose device copy is D; counts to `out` (device memory) */
static int stage_tile(const mdzg_tile* t, mdzg_pview* pv, char* H, char* D, uint32_t* out)
{
    const size_t vsb = 4 * (size_t)pv->vs;
    memset(H, 0, NA_OFF_COORDS);          /* descriptor, counter and cancel flag */
    mdzg_dtile* d = (mdzg_dtile*)H;
    d->out = out;
    d->ncols = t->ncols;
    d->nrows = t->nlines;
    d->col0 = t->col0;
    d->line0 = t->line0;
    d->W = t->view.real_width;
    d->fractal = t->view.fractal;
    d->family = t->view.family;
    d->kclr = pv->kclr;
    d->depth = (int32_t)t->view.depth;
    d->vs = pv->vs;
    d->gprec = pv->v.path == MDZG_PATH_GMP ? pv->nw : 0;
    d->pix_begin = 0;
    if (!pv->vs) {
        mdzg_coords_fp64(pv, &d->xmin, &d->ymax, &d->width, &d->jre, &d->jim);
        return MDZG_OK;
    }
    uint32_t* xs = (uint32_t*)(H + NA_OFF_COORDS);
    uint32_t* ys = (uint32_t*)(H + NA_OFF_COORDS + vsb * (size_t)t->ncols);
    uint32_t* jc = (uint32_t*)(H + NA_OFF_COORDS + vsb * ((size_t)t->ncols + t->nlines));
    d->xs = (const uint32_t*)(D + NA_OFF_COORDS);
    d->ys = (const uint32_t*)(D + NA_OFF_COORDS + vsb * (size_t)t->ncols);
    d->jc = (const uint32_t*)(D + NA_OFF_COORDS + vsb * ((size_t)t->ncols + t->nlines));
    return mdzg_coords_fill(pv, t->line0, t->nlines, t->col0, t->ncols, xs, ys, jc);
}

int mdzg_render_tile_noalloc(mdzg_ctx* c, const mdzg_tile* t, mdzg_stats* st)
{
    const double t0 = mdzg_now_ms();
    if (!c || !tile_ok(t) || c->device < 0 || c->device >= NA_MAX_DEVICES)
        return MDZG_ERR_ARG;
    if (!c->d_scratch || !c->h_scratch)
        return MDZG_ERR_STATE;
    hipSetDevice(c->device);
    /* the counts must be device memory of this device */
    hipPointerAttribute_t pa;
    memset(&pa, 0, sizeof pa);
    if (hipPointerGetAttributes(&pa, t->counts) != hipSuccess ||
        pa.type != hipMemoryTypeDevice || pa.device != c->device) {
        (void)hipGetLastError();
        return MDZG_ERR_ARG;
    }
    mdzg_pview pv;
    int rc = mdzg_view_parse(&t->view, &pv);
    if (rc != MDZG_OK)
        return rc;
    const size_t need = stage_bytes(t, &pv);
    if (need > MDZG_NOALLOC_SCRATCH_BYTES) {
        mdzg_view_clear(&pv);
        return MDZG_ERR_RANGE;
    }
    pthread_mutex_lock(&na_lock[c->device]);
    char* H = (char*)c->h_scratch;
    char* D = (char*)c->d_scratch;
    rc = stage_tile(t, &pv, H, D, t->counts);
    const int kid = pv.kid;
    mdzg_view_clear(&pv);
    if (rc == MDZG_OK) {
        const uint64_t total = (uint64_t)t->ncols * (uint64_t)t->nlines;
        const uint64_t cap = (uint64_t)c->ncu * (uint64_t)c->occ[kid];
        const uint64_t blocks = (total + 255) / 256;
        const int grid = (int)(blocks < cap ? blocks : cap);
        if (hipMemcpyAsync(D, H, need, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipEventRecord(c->na_ev0, c->stream) != hipSuccess ||
            mdzg_dev_launch_persist(kid, (const mdzg_dtile*)D, 1, total,
                                    (unsigned long long*)(D + NA_OFF_CTR),
                                    (const int*)(D + NA_OFF_CTR + 8), grid, c->chunk[kid],
                                    c->stream) != 0 ||
            hipEventRecord(c->na_ev1, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
            hipStreamSynchronize(c->stream);
            rc = MDZG_ERR_DEVICE;
        } else if (st) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->na_ev0, c->na_ev1);
            st->kernel_ms = ms;
            st->wall_ms = mdzg_now_ms() - t0;
            st->pixel_iters = 0.0;
            st->launches = 1;
        }
    }
    pthread_mutex_unlock(&na_lock[c->device]);
    return rc;
}

/* The per-operation probe of the RN arithmetic (measurement only): the
   tile's view must be on the MPFR path at p = 128 or 1024 (the 128- and
   1024-bit legs, whole words); its pixels give the lanes their orbits. */
_Static_assert((int)MDZG_OPP_PRODUCT == (int)MDZG_OPP_MUL &&
               (int)MDZG_OPP_SQUARE == (int)MDZG_OPP_SQR &&
               (int)MDZG_OPP_ADD_SUB == (int)MDZG_OPP_ADD &&
               (int)MDZG_OPP_SQUARES_DIFF == (int)MDZG_OPP_SUBNN &&
               (int)MDZG_OPP_ESCAPE == (int)MDZG_OPP_ESC &&
               (int)MDZG_OPP_ITERATION == (int)MDZG_OPP_STEP &&
               (int)MDZG_OPP_LOOP_ONLY == (int)MDZG_OPP_LOOP, "probe operation numbers");

int mdzg_op_probe(mdzg_ctx* c, const mdzg_tile* t, int op, int warm, int reps, double* kernel_ms)
{
    if (!c || !t || !kernel_ms || op < 0 || op >= MDZG_NOPP || warm < 0 || warm > 1000000 ||
        reps < 1 || reps > 1000000 || t->ncols < 1 || t->nlines < 1 ||
        (long long)t->ncols * t->nlines > (1ll << 24))
        return MDZG_ERR_ARG;
    if (t->view.path != MDZG_PATH_MPFR || (t->view.precision != 128 && t->view.precision != 1024))
        return MDZG_ERR_RANGE;
    hipSetDevice(c->device);
    mdzg_pview pv;
    int rc = mdzg_view_parse(&t->view, &pv);
    if (rc != MDZG_OK)
        return rc;
    const uint32_t npix = (uint32_t)t->ncols * (uint32_t)t->nlines;
    const size_t sb = ALIGN_UP(stage_bytes(t, &pv), 256);
    char* H = (char*)malloc(sb);
    char* D = NULL;
    hipEvent_t e0 = NULL, e1 = NULL;
    if (!H || hipMalloc((void**)&D, sb + 4 * (size_t)npix) != hipSuccess) {
        free(H);
        mdzg_view_clear(&pv);
        return MDZG_ERR_NOMEM;
    }
    rc = stage_tile(t, &pv, H, D, (uint32_t*)(D + sb));
    const int nwords = (int)t->view.precision / 32;
    mdzg_view_clear(&pv);
    float ms = 0.f;
    if (rc == MDZG_OK &&
        (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess ||
         hipMemcpyAsync(D, H, sb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
         hipEventRecord(e0, c->stream) != hipSuccess ||
         mdzg_dev_launch_op_probe(nwords, (const mdzg_dtile*)D, npix, op, warm, reps,
                                  c->stream) != 0 ||
         hipEventRecord(e1, c->stream) != hipSuccess ||
         hipStreamSynchronize(c->stream) != hipSuccess ||
         hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = MDZG_ERR_DEVICE;
    hipStreamSynchronize(c->stream);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    hipFree(D);
    free(H);
    *kernel_ms = ms;
    return rc;
}
