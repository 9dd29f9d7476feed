/*
 * mdzg_host.h -- internal interfaces of the host runtime (C99).
 */
#ifndef MDZG_HOST_H
#define MDZG_HOST_H

#include <gmp.h>
#include <mpfr.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>

#include <hip/hip_runtime_api.h>

#include "../../include/mdzg.h"
#include "../csrc/mdzg_dev.h"

/* kernel table (launch.hip) */
int mdzg_dev_nkernels(void);
int mdzg_dev_occupancy(int kid, int persist);
int mdzg_dev_launch_tile(int kid, const mdzg_dtile* d_tile, uint32_t npix, hipStream_t st);
int mdzg_dev_launch_persist(int kid, const mdzg_dtile* d_tiles, int ntiles, uint64_t total,
                            unsigned long long* d_counter, const int* d_cancel, int grid,
                            int chunk, hipStream_t st);
int mdzg_dev_occupancy_slice(int kid);
int mdzg_dev_rate_probe(int kind, int ncu, int blocks_per_cu, int iters, double* ms);
int mdzg_dev_launch_op_probe(int nwords, const mdzg_dtile* d_tile, uint32_t npix, int op,
                             int warm, int reps, hipStream_t st);
/* launch nodes (launch.hip, k_colour.hip) and their graph replay (graph.hip) */
int mdzg_dev_node_fill(mdzg_gnode* n, uint32_t* p, uint64_t count, uint32_t v);
int mdzg_dev_node_persist(mdzg_gnode* n, int kid, const mdzg_dtile* d_tiles, int ntiles,
                          uint64_t total, unsigned long long* d_counter, const int* d_cancel,
                          int grid, int chunk);
int mdzg_dev_node_slice(mdzg_gnode* n, int kid, const mdzg_dtile* d_tiles, int ntiles,
                        uint64_t total, unsigned long long* d_counter, int grid, int budget,
                        int lane_budget);
int mdzg_dev_node_colour(mdzg_gnode* node, const uint32_t* counts, int real_w, int user_w,
                         int user_h, int aa, const uint32_t* palette, int n, int offset,
                         double sc, int ip, uint32_t depth, uint32_t* rgb,
                         unsigned long long* iters, int max_blocks);
int mdzg_dev_node_launch(const mdzg_gnode* n, hipStream_t st);
typedef struct mdzg_gcache mdzg_gcache;
mdzg_gcache* mdzg_gcache_create(void);
void mdzg_gcache_destroy(mdzg_gcache* gc);
void mdzg_gcache_counts(mdzg_gcache* gc, uint64_t* hits, uint64_t* builds);
int mdzg_graph_run(mdzg_gcache* gc, const mdzg_gnode* root, const mdzg_gnode* par, int npar,
                   const mdzg_gnode* tail, hipStream_t st);
int mdzg_dev_launch_slice(int kid, const mdzg_dtile* d_tiles, int ntiles, uint64_t total,
                          unsigned long long* d_counter, int grid, int budget, int lane_budget,
                          hipStream_t st);

/* k_colour.hip: counts -> RGB (palette_apply / do_anti_aliasing), adding the
   pixel-iterations of the counts read to *iters when iters is not null; at
   most max_blocks blocks of 256 lanes stride over the image */
int mdzg_dev_launch_colour(const uint32_t* counts, int real_w, int user_w, int user_h, int aa,
                           const uint32_t* palette, int n, int offset, double sc, int ip,
                           uint32_t depth, uint32_t* rgb, unsigned long long* iters,
                           int max_blocks, hipStream_t st);

/* An image job: the one tile of the job, user_w*aa x user_h*aa samples,
   coloured on the device after its escape-time kernels. */
typedef struct mdzg_image_req {
    int user_w, user_h, aa;
    const mdzg_palette* pal;
    uint32_t* rgb;              /* user_w x user_h colours out */
    uint32_t* counts;           /* the frame's counts out, or NULL */
} mdzg_image_req;
int mdzg_rt_render_image(mdzg_ctx* c, const mdzg_tile* tile, const mdzg_image_req* img,
                         mdzg_stats* st);
/* the same as an asynchronous job (mdzg_wait / mdzg_job_free); img must stay
   valid until the job is waited for */
int mdzg_rt_submit_image(mdzg_ctx* c, const mdzg_tile* tile, const mdzg_image_req* img,
                         mdzg_job** out);
/* resident blocks per CU of the colouring pass's grid */
#define MDZG_COLOUR_BLOCKS_PER_CU 8
/* validation shared by the colouring entry points */
int mdzg_palette_ok(const mdzg_palette* p);

extern const int mdzg_nwords[MDZG_NCLASS_N];
extern const int mdzg_gmp_limbs[MDZG_NGMP];

/* A parsed view: image_info fields as MPFR values at max(p, 80) bits; on the
   GMP path the coordinates at the coords precision, the Julia c_re at
   max(p, 80) and c_im at 80 bits, as the reference holds them. */
typedef struct mdzg_pview {
    mdzg_view v;              /* caller's scalars (string pointers not kept) */
    long hp;                  /* max(p, 80) */
    mpfr_t xmin, xmax, ymax, width, jre, jim;
    int kid;                  /* kernel id */
    int nw;                   /* words N (RN classes), 0 for FP64; GMP: PREC + 1 limbs */
    int kclr;                 /* 32N - p */
    int vs;                   /* words per stored value, 0 for FP64 */
    int parsed;
} mdzg_pview;

int  mdzg_view_parse(const mdzg_view* v, mdzg_pview* pv);
void mdzg_view_clear(mdzg_pview* pv);

/* kernel id for a view (path, precision); -1 when out of range */
int mdzg_class_of(int path, long precision, int* nw, int* kclr);

/* Per-column / per-row start values of a rectangle, in stored-value format
   (MDZG_VS(nw) words each): xs[ncols], ys[nlines], and jc[2] for Julia. */
int mdzg_coords_fill(const mdzg_pview* pv, int line0, int nlines, int col0, int ncols,
                     uint32_t* xs, uint32_t* ys, uint32_t* jc);

/* dst = src (n counts), returning their pixel-iterations (zero counts
   standing for depth) */
double mdzg_take_counts(uint32_t* dst, const uint32_t* src, size_t n, double depth);

/* FP64 descriptor fields (xmin, ymax, width, Julia c) */
void mdzg_coords_fp64(const mdzg_pview* pv, double* xmin, double* ymax, double* width,
                      double* jre, double* jim);

/* mpf value (at most `limbs` limbs in use) -> GMP-path stored value of a
   bucket of `store` limbs */
void mdzg_store_gf(uint32_t* w, int store, mpf_srcptr x);
/* mpfr (precision <= 32*nw) -> stored value */
void mdzg_store_mpfr(uint32_t* w, int nw, mpfr_srcptr x);
/* long double (x87, 64-bit significand) -> stored value with nw >= 2 */
void mdzg_store_ld(uint32_t* w, int nw, long double x);

/* device context (runtime.c embeds it with its stream and buffer pools) */
struct mdzg_ctx {
    int device;
    int ncu;
    hipStream_t stream;         /* created with the context; jobs use the pool */
    hipStream_t ctl;            /* cancel-flag writes */
    int* d_cancel_zero;         /* a device int that stays 0 */
    int occ[MDZG_NKERNELS];     /* resident blocks per CU (persistent grid) */
    int chunk[MDZG_NKERNELS];   /* iterations between lane refills */
    pthread_mutex_t lock;
    /* mdzg_render_tile_noalloc: the device address of the library's static
       scratch array on this device (noalloc.hip), the host buffer its
       contents are staged in, and the two events that time its launch, all
       set up by mdzg_create */
    void* d_scratch;
    void* h_scratch;
    hipEvent_t na_ev0, na_ev1;
    /* options: graph replay, the line adapter's bands */
    int graphs;
    long long band_pixels;
    int band_inflight;
    mdzg_gcache* gc;            /* graphs of persistent batches */
    /* tests: the launch plan numbered fail_plan (counting from 1, every
       job's plan counts) fails as a device error would, after its copy-in
       was queued (environment variable MDZG_FAILPOINT_PLAN read by
       mdzg_create; 0 = never) */
    long fail_plan;
    long plans;
};

/* noalloc.hip: device address of the static scratch array of the current
   device (MDZG_NOALLOC_SCRATCH_BYTES); 0 on success */
int mdzg_dev_scratch(int device, void** ptr, size_t* bytes);

double mdzg_now_ms(void);

/* the line adapter's band defaults (MDZG_OPT_BAND_PIXELS / _INFLIGHT) */
#define MDZG_DEFAULT_BAND_PIXELS (1 << 18)
#define MDZG_DEFAULT_BAND_INFLIGHT 3

/* directory of the loaded libmdzg.so (tune.c); 0 on success */
int mdzg_lib_dir(char* buf, size_t cap);

#endif
