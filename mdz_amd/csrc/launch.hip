/*
 * launch.hip -- kernel table and launch wrappers called by the C host runtime.
 *
 * Kernel id: 0 = FP64; 1 + 2*i + full for the RN kernels of word count
 * MDZG_NWORDS[i] (full = the tile precision is exactly 32N bits);
 * MDZG_KID_GMP0 + j for the GMP-path kernels of storage bucket j.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "mdzg_dev.h"

extern "C" {
void mdzg_kget_fp64(const void** tile, const void** persist, const void** slice);
void mdzg_kget_rn1(int, const void**, const void**, const void**);
void mdzg_kget_rn2(int, const void**, const void**, const void**);
void mdzg_kget_rn3(int, const void**, const void**, const void**);
void mdzg_kget_rn4(int, const void**, const void**, const void**);
void mdzg_kget_rn6(int, const void**, const void**, const void**);
void mdzg_kget_rn8(int, const void**, const void**, const void**);
void mdzg_kget_rn12(int, const void**, const void**, const void**);
void mdzg_kget_rn16(int, const void**, const void**, const void**);
void mdzg_kget_rn24(int, const void**, const void**, const void**);
void mdzg_kget_rn32(int, const void**, const void**, const void**);
void mdzg_kget_gmp(int bucket, const void**, const void**, const void**);
const void* mdzg_kprobe_rn4(void);
const void* mdzg_kprobe_rn32(void);
}

typedef void (*kget_fn)(int, const void**, const void**, const void**);
static const kget_fn kget_rn[MDZG_NCLASS_N] = {
    mdzg_kget_rn1, mdzg_kget_rn2, mdzg_kget_rn3, mdzg_kget_rn4, mdzg_kget_rn6,
    mdzg_kget_rn8, mdzg_kget_rn12, mdzg_kget_rn16, mdzg_kget_rn24, mdzg_kget_rn32,
};

/* form: 0 = per-tile, 1 = persistent, 2 = resumable slice */
static void lookup(int kid, int form, const void** f)
{
    const void* t = nullptr;
    const void* p = nullptr;
    const void* sl = nullptr;
    if (kid == 0)
        mdzg_kget_fp64(&t, &p, &sl);
    else if (kid >= MDZG_KID_GMP0)
        mdzg_kget_gmp(kid - MDZG_KID_GMP0, &t, &p, &sl);
    else
        kget_rn[(kid - 1) >> 1]((kid - 1) & 1, &t, &p, &sl);
    *f = form == 2 ? sl : (form == 1 ? p : t);
}

/* counts to the UINT_MAX sentinel, queue counters to zero: a kernel, so
   that a graph holds no memset node */
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t* p, uint64_t n, uint32_t v)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        p[i] = v;
}

/* Issue-rate probes for the roofline peak: every lane runs four independent
   chains of one instruction, v_mad_u64_u32 (32x32+64, the multiply-accumulate
   of the RN products) or v_fma_f64, `iters` times.  The chains never stop on
   a data value; the result is stored so nothing is removed. */
__global__ __launch_bounds__(256) void k_probe_mad(uint32_t* out, int iters)
{
    uint64_t a0 = threadIdx.x, a1 = a0 + 1, a2 = a0 + 2, a3 = a0 + 3;
    const uint32_t x = blockIdx.x | 1u, y = threadIdx.x | 3u;
    for (int i = 0; i < iters; ++i) {
        uint64_t c0, c1, c2, c3;
        asm volatile("v_mad_u64_u32 %0, %4, %8, %9, %0\n\t"
                     "v_mad_u64_u32 %1, %5, %8, %9, %1\n\t"
                     "v_mad_u64_u32 %2, %6, %8, %9, %2\n\t"
                     "v_mad_u64_u32 %3, %7, %8, %9, %3"
                     : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "=&s"(c0), "=&s"(c1), "=&s"(c2),
                       "=&s"(c3)
                     : "v"(x), "v"(y));
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(a0 ^ a1 ^ a2 ^ a3);
}

__global__ __launch_bounds__(256) void k_probe_fma(uint32_t* out, int iters)
{
    double a0 = threadIdx.x, a1 = a0 + 1.0, a2 = a0 + 2.0, a3 = a0 + 3.0;
    const double x = 0.999999, y = 1e-9 * (blockIdx.x + 1);
    for (int i = 0; i < iters; ++i) {
        asm volatile("v_fma_f64 %0, %4, %0, %5\n\t"
                     "v_fma_f64 %1, %4, %1, %5\n\t"
                     "v_fma_f64 %2, %4, %2, %5\n\t"
                     "v_fma_f64 %3, %4, %3, %5"
                     : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3)
                     : "v"(x), "v"(y));
    }
    out[blockIdx.x * blockDim.x + threadIdx.x] = (uint32_t)(int)(a0 + a1 + a2 + a3);
}

static void node_init(mdzg_gnode* n, const void* f, unsigned grid, unsigned block)
{
    memset(n, 0, sizeof *n);
    n->func = f;
    n->grid = grid;
    n->block = block;
}

template <class T>
static void node_arg(mdzg_gnode* n, const T& v)
{
    static_assert(sizeof(T) <= MDZG_GNODE_BUF, "argument too large");
    const int o = (n->used + 7) & ~7;
    memcpy((char*)n->buf + o, &v, sizeof v);
    n->off[n->nargs++] = o;
    n->used = o + (int)sizeof v;
}

extern "C" {

int mdzg_dev_nkernels(void) { return MDZG_NKERNELS; }

/* kind 0: v_mad_u64_u32, 1: v_fma_f64.  blocks of 256 lanes (4 waves, one
   per SIMD) per CU x ncu, iters x 4 instructions per lane; *ms the kernel
   time by HIP events on a stream of its own. */
int mdzg_dev_rate_probe(int kind, int ncu, int blocks_per_cu, int iters, double* ms)
{
    if ((kind != 0 && kind != 1) || ncu < 1 || blocks_per_cu < 1 || iters < 1 || !ms)
        return -1;
    const unsigned grid = (unsigned)ncu * (unsigned)blocks_per_cu;
    uint32_t* out = nullptr;
    hipStream_t st = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = -3;
    if (hipMalloc((void**)&out, (size_t)grid * 256 * 4) == hipSuccess &&
        hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess &&
        hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess) {
        const void* f = kind == 0 ? (const void*)k_probe_mad : (const void*)k_probe_fma;
        void* args[] = { (void*)&out, (void*)&iters };
        float t = 0.f;
        if (hipLaunchKernel(f, dim3(grid), dim3(256), args, 0, st) == hipSuccess &&   /* warm-up */
            hipEventRecord(e0, st) == hipSuccess &&
            hipLaunchKernel(f, dim3(grid), dim3(256), args, 0, st) == hipSuccess &&
            hipEventRecord(e1, st) == hipSuccess && hipStreamSynchronize(st) == hipSuccess &&
            hipEventElapsedTime(&t, e0, e1) == hipSuccess) {
            *ms = t;
            rc = 0;
        }
    }
    if (st) hipStreamSynchronize(st);
    if (e0) hipEventDestroy(e0);
    if (e1) hipEventDestroy(e1);
    if (st) hipStreamDestroy(st);
    if (out) hipFree(out);
    return rc;
}

int mdzg_dev_node_fill(mdzg_gnode* n, uint32_t* p, uint64_t count, uint32_t v)
{
    if (!n || !p || count == 0)
        return -1;
    uint64_t g = (count + 255) / 256;
    if (g > 2048)
        g = 2048;
    node_init(n, (const void*)k_fill_u32, (unsigned)g, 256);
    node_arg(n, p);
    node_arg(n, count);
    node_arg(n, v);
    return 0;
}

int mdzg_dev_node_persist(mdzg_gnode* n, int kid, const mdzg_dtile* d_tiles, int ntiles,
                          uint64_t total, unsigned long long* d_counter, const int* d_cancel,
                          int grid, int chunk)
{
    if (!n || kid < 0 || kid >= mdzg_dev_nkernels() || ntiles <= 0 || grid <= 0 || chunk <= 0)
        return -1;
    const void* f;
    lookup(kid, 1, &f);
    node_init(n, f, (unsigned)grid, 256);
    node_arg(n, d_tiles);
    node_arg(n, ntiles);
    node_arg(n, total);
    node_arg(n, d_counter);
    node_arg(n, d_cancel);
    node_arg(n, chunk);
    return 0;
}

int mdzg_dev_node_slice(mdzg_gnode* n, int kid, const mdzg_dtile* d_tiles, int ntiles,
                        uint64_t total, unsigned long long* d_counter, int grid, int budget,
                        int lane_budget)
{
    if (!n || kid < 0 || kid >= mdzg_dev_nkernels() || ntiles <= 0 || grid <= 0 || budget <= 0 ||
        lane_budget <= 0)
        return -1;
    const void* f;
    lookup(kid, 2, &f);
    node_init(n, f, (unsigned)grid, 256);
    node_arg(n, d_tiles);
    node_arg(n, ntiles);
    node_arg(n, total);
    node_arg(n, d_counter);
    node_arg(n, budget);
    node_arg(n, lane_budget);
    return 0;
}

int mdzg_dev_occupancy(int kid, int persist)
{
    if (kid < 0 || kid >= mdzg_dev_nkernels())
        return 0;
    const void* f;
    lookup(kid, persist, &f);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, 0) != hipSuccess)
        return 0;
    return nb;
}

int mdzg_dev_launch_tile(int kid, const mdzg_dtile* d_tile, uint32_t npix, hipStream_t st)
{
    if (kid < 0 || kid >= mdzg_dev_nkernels() || npix == 0)
        return -1;
    const void* f;
    lookup(kid, 0, &f);
    void* args[] = { (void*)&d_tile };
    return hipLaunchKernel(f, dim3((npix + 255u) / 256u), dim3(256), args, 0, st) == hipSuccess
               ? 0 : -3;
}

/* the per-operation probe of the RN kernel of nwords (4 or 32) words, one
   lane per pixel of the tile */
int mdzg_dev_launch_op_probe(int nwords, const mdzg_dtile* d_tile, uint32_t npix, int op,
                             int warm, int reps, hipStream_t st)
{
    const void* f = nwords == 4 ? mdzg_kprobe_rn4() : (nwords == 32 ? mdzg_kprobe_rn32() : nullptr);
    if (!f || npix == 0 || op < 0 || op >= MDZG_NOPP)
        return -1;
    void* args[] = { (void*)&d_tile, (void*)&op, (void*)&warm, (void*)&reps };
    return hipLaunchKernel(f, dim3((npix + 255u) / 256u), dim3(256), args, 0, st) == hipSuccess
               ? 0 : -3;
}

int mdzg_dev_launch_persist(int kid, const mdzg_dtile* d_tiles, int ntiles, uint64_t total,
                            unsigned long long* d_counter, const int* d_cancel, int grid,
                            int chunk, hipStream_t st)
{
    if (kid < 0 || kid >= mdzg_dev_nkernels() || ntiles <= 0 || grid <= 0 || chunk <= 0)
        return -1;
    const void* f;
    lookup(kid, 1, &f);
    void* args[] = { (void*)&d_tiles, (void*)&ntiles, (void*)&total, (void*)&d_counter,
                     (void*)&d_cancel, (void*)&chunk };
    return hipLaunchKernel(f, dim3(grid), dim3(256), args, 0, st) == hipSuccess ? 0 : -3;
}

int mdzg_dev_occupancy_slice(int kid)
{
    if (kid < 0 || kid >= mdzg_dev_nkernels())
        return 0;
    const void* f;
    lookup(kid, 2, &f);
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, f, 256, 0) != hipSuccess)
        return 0;
    return nb;
}

int mdzg_dev_launch_slice(int kid, const mdzg_dtile* d_tiles, int ntiles, uint64_t total,
                          unsigned long long* d_counter, int grid, int budget, int lane_budget,
                          hipStream_t st)
{
    if (kid < 0 || kid >= mdzg_dev_nkernels() || ntiles <= 0 || grid <= 0 || budget <= 0 ||
        lane_budget <= 0)
        return -1;
    const void* f;
    lookup(kid, 2, &f);
    void* args[] = { (void*)&d_tiles, (void*)&ntiles, (void*)&total, (void*)&d_counter,
                     (void*)&budget, (void*)&lane_budget };
    return hipLaunchKernel(f, dim3(grid), dim3(256), args, 0, st) == hipSuccess ? 0 : -3;
}

} /* extern "C" */
