/*
 * mdzg.h -- C ABI of the MI355X (gfx950) escape-time engine for MDZ.
 *
 * The library replaces MDZ's per-pixel escape-time hot path:
 *   the formula loops  frac_X / frac_X_mpfr       (src/frac_mandel.h:12-23,
 *                      src/frac_burning_ship.h:12-23, src/frac_generalized_celtic.h:12-23,
 *                      src/frac_variant.h:12-23)
 *   the line workers   fractal_calculate_line / fractal_mpfr_calculate_line
 *                      (src/fractal.h:42-43, src/fractal.c:29-257)
 * and sits under the line pool of src/render_threads.c (the next_line_cb of
 * src/render_threads.h:70-71) and the command-line render path
 * (render_to_file, src/render.c:10-105).
 *
 * Results are bit-identical to the reference loops: the MPFR path rounds every
 * operation to nearest-even at the view's precision exactly as MPFR_RNDN does,
 * the long double path equals the x87 loop (round-to-nearest at 64 bits), and
 * the double path equals the same C loop evaluated in IEEE double.
 *
 * Plain C types only: views are described by the image_info fields the line
 * workers read, coordinates as MPFR-parsable strings (decimal or the hex-float
 * form mpfr_asprintf("%Ra") prints, which is exact).
 *
 * Return values: >= 0 success, negative MDZG_ERR_* codes.  Every entry point
 * is safe to call from several host threads on one context.
 */
#ifndef MDZG_H
#define MDZG_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDZG_ABI_VERSION 1

enum {
    MDZG_OK = 0,
    MDZG_ERR_ARG = -1,      /* invalid argument                         */
    MDZG_ERR_NOMEM = -2,    /* host or device allocation failed         */
    MDZG_ERR_DEVICE = -3,   /* HIP runtime error                        */
    MDZG_ERR_STOPPED = -4,  /* render stopped before completion         */
    MDZG_ERR_BUSY = -5,     /* admission refused (budget exhausted)     */
    MDZG_ERR_PARSE = -6,    /* a coordinate string did not parse        */
    MDZG_ERR_RANGE = -7,    /* precision or size beyond what is built   */
    MDZG_ERR_STATE = -8     /* call out of sequence                     */
};

/* fractal types, same values as the enum of src/fractal.h:17-26 */
enum { MDZG_MANDELBROT = 0, MDZG_BURNING_SHIP = 1, MDZG_GENERALIZED_CELTIC = 2,
       MDZG_VARIANT = 3 };
/* families, same values as src/fractal.h:8-15 */
enum { MDZG_FAMILY_MANDEL = 0, MDZG_FAMILY_JULIA = 1 };
/* arithmetic, the choice image_info_set_multi_prec makes (src/image_info.c:226-249):
   MPFR = fractal_mpfr_calculate_line, LD = fractal_calculate_line (long double),
   DOUBLE = the same loop in IEEE double (the FP64 path),
   GMP = fractal_gmp_calculate_line (src/fractal.c:260-397, GMP's truncating mpf
   floats; multi-precision without rounding) */
enum { MDZG_PATH_MPFR = 0, MDZG_PATH_LD = 1, MDZG_PATH_DOUBLE = 2, MDZG_PATH_GMP = 3 };

#define MDZG_MAX_PRECISION 1024   /* largest p of the MPFR and GMP paths built for the device */
#define MDZG_MIN_PRECISION 2      /* smallest p of the MPFR and GMP paths                  */

/* The image_info fields the line workers read (src/image_info.h:63-122). */
typedef struct mdzg_view {
    int fractal;            /* image_info.fractal                                 */
    int family;             /* image_info.family                                  */
    int path;               /* MDZG_PATH_*                                        */
    int real_width;         /* image_info.real_width  (user width x aa)           */
    int real_height;        /* image_info.real_height                             */
    long precision;         /* image_info.precision: p of the MPFR path           */
    long long depth;        /* image_info.depth (1 .. INT32_MAX)                  */
    const char* xmin;       /* image_info.xmin, .xmax, .ymax, .width (held at     */
    const char* xmax;       /*   max(p, 80) bits, as image_info_set_precision     */
    const char* ymax;       /*   keeps them, src/image_info.c:252-280).  GMP path: */
    const char* width;      /*   the coords' own values (mdzg_rect_from_center_coords) */
    const char* julia_re;   /* image_info.u.julia.c_re / c_im (Julia family only) */
    const char* julia_im;   /*   GMP path: read as MDZ holds them, c_re at max(p,80)
                                 bits, c_im at 80 bits (src/image_info.c:63-64, :271-272) */
} mdzg_view;

/* A rectangle of a view: rows [line0, line0+nlines) x columns [col0, col0+ncols),
   counts written row-major (ncols per row). */
typedef struct mdzg_tile {
    mdzg_view view;
    int line0, nlines;
    int col0, ncols;
    uint32_t* counts;       /* host memory, nlines*ncols                          */
} mdzg_tile;

/* Timings and work of the last batch (device clock, HIP events). */
typedef struct mdzg_stats {
    double kernel_ms;       /* first launch start to last launch end on the device */
    double wall_ms;         /* host wall time of the call, staging and copies incl. */
    double pixel_iters;     /* sum over pixels of (count ? count : depth)          */
    int launches;           /* kernel launches made                               */
} mdzg_stats;

typedef struct mdzg_ctx mdzg_ctx;

int         mdzg_abi_version(void);
const char* mdzg_strerror(int code);

/* Open device `device` (HIP ordinal).  One context per process per device. */
int  mdzg_create(int device, mdzg_ctx** out);
void mdzg_destroy(mdzg_ctx* ctx);

/* Rect of a centre/size view as the command-line path computes it
   (coords_set_precision src/coords.c:204-231, coords_center_to_rect :265-302,
   coords_get_rect :305-313).  Writes img->xmin/xmax/ymax/width as hex-float
   strings at max(precision, 80) bits into buffers of `cap` bytes. */
int mdzg_rect_from_center(const char* cx, const char* cy, const char* size,
                          long precision, int user_w, int user_h,
                          char* xmin, char* xmax, char* ymax, char* width, int cap);

/* The same rect as coords holds it, at the coords precision (max(p, 80) raised
   to whole GMP limbs, src/coords.c:204-231): what coords_get_rect_gmp converts
   for the GMP path (src/coords.c:317-324, src/render.c:36-37). */
int mdzg_rect_from_center_coords(const char* cx, const char* cy, const char* size,
                                 long precision, int user_w, int user_h,
                                 char* xmin, char* xmax, char* ymax, char* width, int cap);

/* ---- colouring -----------------------------------------------------------
   A palette as MDZ's colouring reads it: the colours (palette[], RGB(r, g, b)
   = r | g << 8 | b << 16, src/palette.h:16), their count pal_indexes, the
   rotation pal_offset, and image_info's colour_scale and palette_ip. */
typedef struct mdzg_palette {
    uint32_t colours[256];
    int n;                      /* pal_indexes, 2..256 (MDZ indexes modulo n - 1) */
    int offset;                 /* pal_offset, >= 0                               */
    double colour_scale;        /* > 0                                            */
    int interpolate;            /* palette_ip                                     */
} mdzg_palette;

/* A palette file (.map): palette_load / palette_read (src/palette.c:167-217),
   up to 256 lines of "r g b".  Sets colours and n; keeps the other fields. */
int mdzg_palette_load(const char* path, mdzg_palette* pal);

/* Colour a frame of counts, as render_to_file does row by row: palette_apply
   (src/palette.c:406-463) when aa == 1, do_anti_aliasing (src/render.c:108-157)
   when aa > 1.  counts: (user_w * aa) x (user_h * aa) samples, MDZ's raw_data;
   rgb: user_w x user_h colours, MDZ's rgb_data.  On the device; stats (may be
   NULL) times the colouring kernel alone. */
int mdzg_colour(mdzg_ctx* ctx, const uint32_t* counts, int user_w, int user_h, int aa,
                const mdzg_palette* pal, uint32_t* rgb, mdzg_stats* stats);

/* Render a whole view (real_width x real_height = user size x aa) and colour
   it on the device in one job: the counts stay in device memory and only the
   colours come back (and the counts, when counts is not NULL).  stats counts
   the frame's pixel-iterations as mdzg_render_tile does. */
int mdzg_render_image(mdzg_ctx* ctx, const mdzg_view* view, int aa, const mdzg_palette* pal,
                      uint32_t* rgb, uint32_t* counts, mdzg_stats* stats);

/* save_png_file (src/my_png.c:28-109): an 8-bit RGB, non-interlaced PNG of
   w x h rgb_data colours (the byte above blue is dropped, as png_set_filler
   does).  MDZG_ERR_ARG if the file cannot be written. */
int mdzg_png_write(const char* path, const uint32_t* rgb, int w, int h);

/* ---- scene files --------------------------------------------------------
   A fractal settings file (.mdz) as MDZ's command line reads its settings
   section (image_info_load_settings, src/image_info.c:456-721, with
   src/mdzfileio.c and src/setting.c). */
typedef struct mdzg_scene {
    int version_major, version_minor, version_revision;   /* header version    */
    int family, fractal;            /* MDZG_FAMILY_*, MDZG_MANDELBROT..VARIANT    */
    long long depth;
    double aspect;                  /* width / height of the saved image         */
    double colour_scale;
    int colour_interpolate;
    int multi_precision, rounding;  /* the file's flags ...                       */
    int path;                       /* ... and the MDZG_PATH_* they select        */
    long precision;                 /* >= 80                                      */
    int centre;                     /* 1: c0..c2 = cx, cy, size; 0: xmin, xmax, ymax */
    char* c0;                       /* the file's decimal strings                 */
    char* c1;
    char* c2;
    char* julia_re;                 /* Julia family only                          */
    char* julia_im;
    int palette_offset;
    int palette_n;                  /* colours of the palette section: its "data" */
    uint32_t palette[256];          /*   lines or the "file" it names; 0 if none or unreadable */
} mdzg_scene;

int  mdzg_scene_load(const char* path, mdzg_scene* out);   /* MDZG_ERR_PARSE on a bad file */
void mdzg_scene_free(mdzg_scene* s);

/* The scene's palette (its palette section, palette-offset, colour-scale and
   colour-interpolate).  MDZG_ERR_PARSE when the file has no usable palette,
   a file MDZ's command line refuses (image_info_load_all, src/image_info.c:769-783). */
int mdzg_scene_palette(const mdzg_scene* s, mdzg_palette* pal);

/* The image size MDZ's command line renders a scene at (main.c init_misc):
   *w, *h of 0 are filled in -- both 0 gives 480 x (int)(480 / aspect); one 0
   follows the aspect 480 / (int)(480 / aspect) that loading stored. */
int mdzg_scene_dims(const mdzg_scene* s, int* w, int* h);

/* The rect (hex strings, as mdzg_rect_from_center gives them) MDZ's command
   line renders a scene with at user_w x user_h.  Centre form: coords_to +
   coords_size, then coords_center_to_rect -- mdzg_rect_from_center
   (mdzg_rect_from_center_coords on the GMP path).  Corner form:
   coords_set_rect + coords_rect_to_center at the aspect of the 480-wide load
   (src/image_info.c:666-669, src/coords.c:327-333, :234-252); then init_misc's
   coords_set puts coords' stale size, 4.0, back into the width (wide output)
   or height (tall), and coords_center_to_rect (:265-313) centres that on the
   file's rect.  flags = MDZG_SCENE_FIT_RECT instead keeps the file's width
   and refits its rect at the output aspect (no counterpart in the
   reference's command line). */
enum { MDZG_SCENE_FIT_RECT = 1 };
int mdzg_scene_rect(const mdzg_scene* s, int user_w, int user_h, int flags,
                    char* xmin, char* xmax, char* ymax, char* width, int cap);

/* Synchronous render of one tile: what fractal_*_calculate_line computes for
   those rows and columns (src/fractal.c:29-257), in one persistent launch. */
int mdzg_render_tile(mdzg_ctx* ctx, const mdzg_tile* tile, mdzg_stats* stats);

/* ---- render without allocating (no counterpart in the reference) --------
   For callers that cannot allocate device memory: code under HIP graph
   capture, tight serving loops, harnesses that forbid hipMalloc.  The same
   render as mdzg_render_tile, except that tile->counts is DEVICE memory of
   the context's device (nlines * ncols uint32, written in place) and every
   other device buffer of the render -- the tile descriptor, the queue
   counter, the cancel flag and the start values -- lives in a static device
   array of MDZG_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends.
   The call makes no hipMalloc, hipHostMalloc or hipFree; it is synchronous
   on the context's stream.  MDZG_ERR_RANGE when the tile's start values do
   not fit the array: (ncols + nlines + 2) values of 16 * ceil((N + 2) / 4)
   bytes with N = ceil(p / 32) words (MPFR and long double paths) or
   16 * ceil((2L + 3) / 4) bytes with L the GMP limb bucket; FP64 tiles need
   none.  MDZG_ERR_ARG when tile->counts is not device memory of the
   context's device.  Every call on a GPU shares that GPU's array, which
   keeps the last call's contents: calls are serialised within a process
   (calls that overlap wait for each other).  stats->pixel_iters is 0: the
   counts are not read back. */
#define MDZG_NOALLOC_SCRATCH_BYTES (8u << 20)
int mdzg_render_tile_noalloc(mdzg_ctx* ctx, const mdzg_tile* tile, mdzg_stats* stats);

/* Batch modes */
enum { MDZG_BATCH_PERSISTENT = 0,   /* one persistent launch per arithmetic class,
                                       draining a device queue of all tiles' pixels */
       MDZG_BATCH_PER_TILE = 1 };   /* one plain launch per tile (baseline form) */

/* Render many tiles (uneven size, precision, depth, formula).  Counts are
   identical in both modes. */
int mdzg_render_batch(mdzg_ctx* ctx, const mdzg_tile* tiles, int ntiles, int mode,
                      mdzg_stats* stats);

/* ---- tuning (no counterpart in the reference) ---------------------------
   Kernel ids: mdzg_kernel_id(path, precision) gives the kernel a view uses
   (0 = FP64; the RN kernels by word count and exact/ragged precision).
   blocks_per_cu: resident 256-thread blocks per CU the persistent grid is
   sized for (<= the kernel's occupancy limit, reported as max_blocks_per_cu);
   chunk: iterations a lane runs between queue refills.  0 keeps a value. */
int mdzg_kernel_id(int path, long precision);
int mdzg_tune_set(mdzg_ctx* ctx, int kernel_id, int blocks_per_cu, int chunk);
int mdzg_tune_get(mdzg_ctx* ctx, int kernel_id, int* blocks_per_cu, int* chunk,
                  int* max_blocks_per_cu);

/* ---- options (no counterpart in the reference) --------------------------
   MDZG_OPT_GRAPHS        1 (default): a persistent batch of two or more
                          arithmetic classes (its sentinel fill, its
                          concurrent class launches and an image job's
                          colouring pass) and each serving slice of two or
                          more classes (counter reset and class launches) run
                          as one replay of a cached HIP graph; plans of one
                          class are launched directly.  0: every plan is
                          launched one by one, the classes forked onto streams.
   MDZG_OPT_BAND_PIXELS   the line adapter's band size in pixels (a band is
                          whole rows, at least this many pixels); 1..2^30.
   MDZG_OPT_BAND_INFLIGHT bands the line adapter keeps in flight; 1..16.
   Options set before mdzg_frame_open or mdzg_server_create apply to it. */
enum { MDZG_OPT_GRAPHS = 1, MDZG_OPT_BAND_PIXELS = 2, MDZG_OPT_BAND_INFLIGHT = 3 };
int mdzg_option_set(mdzg_ctx* ctx, int option, long long value);
int mdzg_option_get(mdzg_ctx* ctx, int option, long long* value);

/* ---- autotuning (no counterpart in the reference) -----------------------
   mdzg_autotune renders `probes` (1 to 8 tiles of the kernel class to tune,
   e.g. one whose pixels escape at many iterations and one whose pixels all
   run to depth) at every candidate pair of blocks per CU (1 .. the kernel's
   occupancy limit) and chunk (iterations between refills: 1, 2, 4, .. 32 on
   the RN and GMP paths, 64 .. 1024 on FP64), `reps` times each, timing the
   kernel with HIP events.  A pair's cost is the mean over the probes of its
   time over the time of the pair in place; the cheapest pair is kept in the
   context (mdzg_tune_set) when its cost is below 0.98, beyond the probes'
   run-to-run spread.  best_ms and default_ms sum the probes' times.
   mdzg_tune_save writes the context's table (every kernel's pair and the
   options above) as text; mdzg_tune_load reads one.  mdzg_create loads the
   table named by the environment variable MDZG_TUNE_TABLE, else the file
   tune_gfx950.txt beside the library, when it exists. */
typedef struct mdzg_tune_result {
    int kernel_id;
    int blocks_per_cu, chunk;         /* the pair kept                         */
    double best_ms, default_ms;       /* its probes' kernel time, the previous pair's */
    int candidates;                   /* pairs measured                        */
} mdzg_tune_result;
int mdzg_autotune(mdzg_ctx* ctx, const mdzg_tile* probes, int nprobes, int reps,
                  mdzg_tune_result* out);
int mdzg_tune_save(mdzg_ctx* ctx, const char* path);
int mdzg_tune_load(mdzg_ctx* ctx, const char* path);

/* ---- issue-rate probe (no counterpart in the reference) ---------------
   The roofline peak of the RN kernels: *ops_per_s of v_mad_u64_u32 (kind 0,
   32x32+64 multiply-accumulates, the products' instruction) or v_fma_f64
   (kind 1) lane operations, measured with blocks_per_cu resident 256-lane
   blocks per CU, four independent chains per lane.  Allocates a small
   buffer of its own. */
int mdzg_rate_probe(mdzg_ctx* ctx, int kind, int blocks_per_cu, double* ops_per_s);

/* ---- per-operation probe (no counterpart in the reference) -------------
   Measurement only.  One lane per pixel of the tile (MPFR path, p = 128 or
   1024: the 128- and 1024-bit legs); each lane runs `warm` iterations of
   the escape loop from its pixel, then repeats one RN operation `reps`
   times in a dependent chain (op: 0 loop only, 1 product, 2 square,
   3 addition of c_re and its subtraction (two a rep), 4 difference of the
   squares, 5 escape test, 6 whole iteration).  *kernel_ms is the launch's
   time; instruction counters of the launch divided by its waves and reps
   give the operation's cost.  tile->counts is not used (the probe writes a
   buffer of its own).  MDZG_ERR_RANGE for another path or precision. */
enum { MDZG_OPP_LOOP_ONLY = 0, MDZG_OPP_PRODUCT = 1, MDZG_OPP_SQUARE = 2,
       MDZG_OPP_ADD_SUB = 3, MDZG_OPP_SQUARES_DIFF = 4, MDZG_OPP_ESCAPE = 5,
       MDZG_OPP_ITERATION = 6 };
int mdzg_op_probe(mdzg_ctx* ctx, const mdzg_tile* tile, int op, int warm, int reps,
                  double* kernel_ms);

/* ---- asynchronous form (submit / poll / wait / cancel) ---------------- */
typedef struct mdzg_job mdzg_job;
int  mdzg_submit(mdzg_ctx* ctx, const mdzg_tile* tiles, int ntiles, mdzg_job** out);
int  mdzg_poll(mdzg_job* job);      /* 1 done, 0 running, < 0 error             */
int  mdzg_wait(mdzg_job* job, mdzg_stats* stats); /* MDZG_OK or MDZG_ERR_STOPPED */
int  mdzg_cancel(mdzg_job* job);    /* device-side stop; untouched counts keep
                                       their previous contents                  */
void mdzg_job_free(mdzg_job* job);

/* ---- serving: continuous batching of tile requests ----------------------
   No counterpart in the reference (TODO:15-16 asks for a rendering daemon).
   A scheduler thread admits requests oldest-first while the HBM budget and
   each arithmetic class's resident-pixel budget allow, then runs the device
   in slices: every unfinished pixel of every admitted request resumes from
   its orbit state kept in HBM and runs up to slice_iters more iterations.
   Requests arriving meanwhile join at the next slice.  A request's counts
   are identical alone or in any mix.  The tile's view strings and counts
   buffer must stay valid until the request is done. */
enum { MDZG_SERVE_BATCHED = 0,       /* continuous batching (slices)          */
       MDZG_SERVE_PER_REQUEST = 1 }; /* baseline: one launch per request      */

typedef struct mdzg_server_cfg {
    int mode;                 /* MDZG_SERVE_*                                    */
    size_t hbm_budget;        /* device bytes for resident requests (0: 8 GiB)   */
    double occupancy_budget;  /* resident pixels per class, in units of the lanes
                                 that class's slice kernel keeps resident (0: 4) */
    int slice_iters;          /* iterations per pixel per slice (0: by class)    */
    int max_requests;         /* requests resident at once (0: 4096)             */
} mdzg_server_cfg;

typedef struct mdzg_req_info {
    int status;               /* 0 pending, 1 running, 2 done, < 0 error         */
    double submit_ms, admit_ms, done_ms;   /* host monotonic clock               */
    double latency_ms;        /* done - submit                                   */
    double pixel_iters;
} mdzg_req_info;

typedef struct mdzg_serve_stats {
    uint64_t submitted, admitted, completed, slices, active, pending;
    double pixel_iters;
    size_t hbm_in_use, hbm_peak;
} mdzg_serve_stats;

typedef struct mdzg_server mdzg_server;
int  mdzg_server_create(mdzg_ctx* ctx, const mdzg_server_cfg* cfg, mdzg_server** out);
int  mdzg_server_submit(mdzg_server* srv, const mdzg_tile* tile, uint64_t* id);
/* An image request: the tile (nlines and ncols multiples of aa) is rendered
   like any request and, when done, coloured on the device as mdzg_colour
   colours (the palette is copied).  rgb receives (ncols/aa) x (nlines/aa)
   colours; tile->counts receives the counts unless it is NULL. */
int  mdzg_server_submit_image(mdzg_server* srv, const mdzg_tile* tile, int aa,
                              const mdzg_palette* pal, uint32_t* rgb, uint64_t* id);
int  mdzg_server_poll(mdzg_server* srv, uint64_t id);     /* 1 done, 0 not yet */
int  mdzg_server_wait(mdzg_server* srv, uint64_t id, mdzg_req_info* info);
int  mdzg_server_info(mdzg_server* srv, uint64_t id, mdzg_req_info* info);
int  mdzg_server_stats(mdzg_server* srv, mdzg_serve_stats* stats);
void mdzg_server_destroy(mdzg_server* srv);   /* finishes admitted and pending work */

/* ---- line-callback adapter ---------------------------------------------
   A frame holds one view and the caller's raw buffer (image_info.raw_data,
   real_width*real_height ints, src/image_info.c:136-137).  The frame is
   rendered on the device in row bands; mdzg_frame_calculate_line has the
   shape and return convention of fractal_mpfr_calculate_line (src/fractal.h:31-44):
   it fills row `line` of raw and returns 1, or returns 0 once the frame was
   stopped (rows never rendered keep whatever raw held, e.g. the UINT_MAX
   sentinel of src/image_info.c:214-215).  Safe to call from several worker
   threads at once with distinct lines, as src/render_threads.c:376 does. */
typedef struct mdzg_frame mdzg_frame;
int  mdzg_frame_open(mdzg_ctx* ctx, const mdzg_view* view, int32_t* raw, mdzg_frame** out);
int  mdzg_frame_calculate_line(mdzg_frame* frame, int line);
void mdzg_frame_stop(mdzg_frame* frame);    /* the RT_STOP of src/render_threads.c:409-418 */
int  mdzg_frame_close(mdzg_frame* frame, mdzg_stats* stats);

#ifdef __cplusplus
}
#endif
#endif
