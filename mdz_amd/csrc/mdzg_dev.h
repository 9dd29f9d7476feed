/*
 * mdzg_dev.h -- layout shared by the host runtime (C) and the gfx950 kernels.
 *
 * A device tile is one rectangle of one view.  Pixels of all tiles of one
 * arithmetic class form a flat index space [0, total): tile t owns
 * [pix_begin, pix_begin + nrows*ncols), row-major inside the tile, so the
 * lanes of a wave walk along a row (neighbouring c, similar escape times).
 *
 * Stored multi-word values (mp<N> of mp_rn.h) take MDZG_VS(N) 32-bit words:
 *   w[0..N-1] significand (least significant first), w[N] exponent (int32),
 *   w[N+1] sign, padded to a multiple of 4 words for 16-byte loads.  Values of
 *   the GMP path (gf<L> of mpf_gmp.h) take MDZG_GF_VS(L) words.
 */
#ifndef MDZG_DEV_H
#define MDZG_DEV_H

#include <stdint.h>

#define MDZG_VS(N) ((((N) + 2) + 3) & ~3)

/* arithmetic classes: FP64, the RN kernels for N words (N in MDZG_NWORDS),
   and the GMP-path kernels for values of up to MDZG_GMP_LIMBS 64-bit limbs */
#define MDZG_NCLASS_N 10
#define MDZG_NWORDS_INIT { 1, 2, 3, 4, 6, 8, 12, 16, 24, 32 }
#define MDZG_CLASS_FP64 (-1)
#define MDZG_NGMP 4
#define MDZG_GMP_LIMBS_INIT { 4, 6, 10, 18 }
/* kernel ids: 0 FP64, 1 + 2i + full for RN class i, 1 + 2 MDZG_NCLASS_N + j for
   GMP bucket j */
#define MDZG_NKERNELS (1 + 2 * MDZG_NCLASS_N + MDZG_NGMP)
#define MDZG_KID_GMP0 (1 + 2 * MDZG_NCLASS_N)

/* operations of the RN per-operation probe (k_rn_op_probe) */
enum { MDZG_OPP_LOOP = 0, MDZG_OPP_MUL = 1, MDZG_OPP_SQR = 2, MDZG_OPP_ADD = 3,
       MDZG_OPP_SUBNN = 4, MDZG_OPP_ESC = 5, MDZG_OPP_STEP = 6, MDZG_NOPP = 7 };
/* GMP-path stored value of L limbs: 2L words (limbs from the most
   significant, low word first), size, exponent (limbs), sign; padded to 4 */
#define MDZG_GF_VS(L) (((2 * (L) + 3) + 3) & ~3)

typedef struct mdzg_dtile {
    uint32_t* out;              /* counts: nrows x ncols, row-major                  */
    const uint32_t* xs;         /* RN: ncols stored values (z0.re / c.re per column) */
    const uint32_t* ys;         /* RN: nrows stored values (z0.im / c.im per row)    */
    const uint32_t* jc;         /* RN Julia: c.re, c.im (2 stored values) or 0       */
    double xmin, ymax, width;   /* FP64: view rect, width = xmax - xmin              */
    double jre, jim;            /* FP64 Julia c                                      */
    uint64_t pix_begin;         /* first flat pixel index of this tile in its class  */
    int32_t ncols, nrows;
    int32_t col0, line0;        /* FP64: origin of the rectangle inside the view     */
    int32_t W;                  /* FP64: view real_width (the coordinate divisor)    */
    int32_t fractal, family;
    int32_t kclr;               /* RN: 32N - p bits cleared by rounding              */
    int32_t depth;              /* max iterations (image_info.depth)                 */
    int32_t vs;                 /* RN, GMP: words per stored value                   */
    int32_t gprec;              /* GMP: PREC + 1, the limbs a value may hold         */
    /* resumable slices (serving): per-pixel orbit state between slices */
    uint32_t* state;            /* z, z^2 per pixel: RN 4 stored values, FP64 4 doubles */
    int32_t* wzs;               /* iterations done per pixel (0 = not started)       */
    unsigned int* remaining;    /* unfinished pixels of this request                 */
} mdzg_dtile;

/* One kernel launch as data: the kernel, its grid, and its arguments packed
   in buf at off[i] (each aligned to 8 bytes).  The runtime builds launches
   as nodes and either launches them at once or replays them as the nodes of
   a cached HIP graph (graph.hip). */
#define MDZG_GNODE_ARGS 8
#define MDZG_GNODE_BUF 160
typedef struct mdzg_gnode {
    const void* func;
    unsigned grid, block, shmem;
    int nargs;
    int used;                   /* bytes of buf in use */
    int off[MDZG_GNODE_ARGS];
    uint64_t buf[MDZG_GNODE_BUF / 8];
} mdzg_gnode;

#endif
