/*
 * k_colour.hip -- counts to RGB on the device: MDZ's palette_apply
 * (src/palette.c:406-463, one sample per pixel) and do_anti_aliasing
 * (src/render.c:108-157, the box average of aa x aa samples coloured by
 * get_pixel_colour, src/palette.c:466-502), in one pass over the counts.
 *
 * The colour arithmetic is MDZ's, operation for operation, in IEEE double
 * (the file is built with -ffp-contract=off, so no step is fused), with the
 * float-to-integer conversions as x86-64 performs them:
 *   (guint32)v  -> the low 32 bits of the 64-bit truncation (0 from 2^63 up);
 *   (int)v      -> the 32-bit truncation, which palette_apply's interpolating
 *                  path uses for ceil(v) and floor(v).  Below 2^31 it equals
 *                  (guint32)v; from 2^31 up MDZ's palette index is negative
 *                  (undefined), and the kernel gives get_pixel_colour's
 *                  colour, which uses (guint32)v throughout;
 *   (guint8)v   -> the low byte of the 32-bit truncation (v is in [0, 256)).
 * A colour is RGB(r, g, b) = r | g << 8 | b << 16 (src/palette.h:16), the
 * layout of image_info.rgb_data.
 *
 * The pass reads every count anyway, so it also sums the pixel-iterations of
 * the frame (count, or depth for a pixel that never escaped) into a 64-bit
 * device counter: the escape-time counts need not leave HBM.
 *
 * HBM-bound: 4 aa^2 bytes read and 4 bytes written per output pixel.  The
 * grid is sized to what stays resident and strides over the image, so each
 * block stages the palette (<= 256 colours) in LDS once; aa = 1 moves 4
 * pixels per lane and aa = 2 two output pixels per lane with 16-byte loads.
 * Palette indexes are reduced modulo n - 1 by Lemire's multiply-based
 * remainder (Lemire, Kaser, Kurz, "Faster remainder by direct computation",
 * 2019), exact for every 32-bit index.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "colour_math.h"
#include "mdzg_dev.h"

using mdzg::colour::colour_apply;
using mdzg::colour::colour_get;

struct colour_args {
    const uint32_t* counts;      /* real_w x real_h, row-major                  */
    const uint32_t* palette;     /* n colours                                   */
    uint32_t* rgb;               /* user_w x user_h                             */
    unsigned long long* iters;   /* pixel-iteration sum (added to), may be null */
    double sc;                   /* image_info.colour_scale                     */
    uint64_t mmagic;             /* remainder constant for n - 1                */
    int real_w, user_w, user_h, aa;
    int n;                       /* pal_indexes, 2..256                         */
    int offset;                  /* pal_offset                                  */
    int ip;                      /* image_info.palette_ip                       */
    uint32_t depth;
};

struct pal_ctx {
    mdzg::colour::palette_view P;
    double sc;
    int offset;
};

/* IP: interpolating palette; COUNT: sum the pixel-iterations */
template <bool IP, bool COUNT>
__device__ __forceinline__ uint32_t apply1(const pal_ctx& p, uint32_t raw, uint64_t& it,
                                           uint32_t depth)
{
    if (COUNT)
        it += raw ? raw : depth;
    return colour_apply<IP>(p.P, raw, p.sc, p.offset);
}

/* one sample of the anti-alias box, its channels added to r, g, b */
template <bool IP, bool COUNT>
__device__ __forceinline__ void sample(const pal_ctx& p, uint32_t raw, uint32_t& r, uint32_t& g,
                                       uint32_t& b, uint64_t& it, uint32_t depth)
{
    if (COUNT)
        it += raw ? raw : depth;
    const uint32_t c = colour_get<IP>(p.P, raw, p.sc, p.offset);
    r += c & 0xffu;
    g += (c >> 8) & 0xffu;
    b += (c >> 16) & 0xffu;
}

__device__ __forceinline__ uint32_t pack_avg(uint32_t r, uint32_t g, uint32_t b, uint32_t aaaa)
{
    return (r / aaaa) | ((g / aaaa) << 8) | ((b / aaaa) << 16);
}

/* FORM 1: aa = 1, counts contiguous (real_w = user_w), 4 pixels per lane;
   FORM 2: aa = 2, user_w even, 2 output pixels (4 x 2 samples) per lane;
   FORM 0: any aa, one output pixel per lane */
template <int FORM, bool IP, bool COUNT>
__global__ __launch_bounds__(256) void k_colour(colour_args a)
{
    __shared__ uint32_t lpal[256];
    __shared__ unsigned long long lsum[4];
    for (int i = threadIdx.x; i < a.n; i += blockDim.x)
        lpal[i] = a.palette[i];
    __syncthreads();

    pal_ctx p;
    p.P.pal = lpal;
    p.P.M = a.mmagic;
    p.P.m = (uint32_t)(a.n - 1);
    p.sc = a.sc;
    p.offset = a.offset;
    const uint32_t depth = a.depth;
    const uint32_t npix = (uint32_t)a.user_w * (uint32_t)a.user_h;
    const uint32_t stride = gridDim.x * blockDim.x;
    const uint32_t t0 = blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t it = 0;

    if (FORM == 1) {
        const uint32_t nq = npix >> 2;
        const uint4* src = (const uint4*)a.counts;
        uint4* dst = (uint4*)a.rgb;
        for (uint32_t q = t0; q < nq; q += stride) {
            const uint4 c = src[q];
            uint4 o;
            o.x = apply1<IP, COUNT>(p, c.x, it, depth);
            o.y = apply1<IP, COUNT>(p, c.y, it, depth);
            o.z = apply1<IP, COUNT>(p, c.z, it, depth);
            o.w = apply1<IP, COUNT>(p, c.w, it, depth);
            dst[q] = o;
        }
        const uint32_t k = (nq << 2) + t0;           /* the last npix % 4 pixels */
        if (t0 < (npix & 3u))
            a.rgb[k] = apply1<IP, COUNT>(p, a.counts[k], it, depth);
    } else if (FORM == 2) {
        const uint32_t hw = (uint32_t)a.user_w >> 1;
        const uint32_t npair = npix >> 1;
        for (uint32_t q = t0; q < npair; q += stride) {
            const uint32_t y = q / hw;
            const uint32_t x = (q - y * hw) << 1;
            const uint32_t* r0 = a.counts + (size_t)(2 * y) * (size_t)a.real_w + 2 * x;
            const uint4 u = *(const uint4*)r0;
            const uint4 v = *(const uint4*)(r0 + a.real_w);
            uint32_t r = 0, g = 0, b = 0;
            sample<IP, COUNT>(p, u.x, r, g, b, it, depth);
            sample<IP, COUNT>(p, u.y, r, g, b, it, depth);
            sample<IP, COUNT>(p, v.x, r, g, b, it, depth);
            sample<IP, COUNT>(p, v.y, r, g, b, it, depth);
            uint2 o;
            o.x = pack_avg(r, g, b, 4u);
            r = g = b = 0;
            sample<IP, COUNT>(p, u.z, r, g, b, it, depth);
            sample<IP, COUNT>(p, u.w, r, g, b, it, depth);
            sample<IP, COUNT>(p, v.z, r, g, b, it, depth);
            sample<IP, COUNT>(p, v.w, r, g, b, it, depth);
            o.y = pack_avg(r, g, b, 4u);
            *(uint2*)(a.rgb + (size_t)y * (size_t)a.user_w + x) = o;
        }
    } else {
        const int aa = a.aa;
        const uint32_t aaaa = (uint32_t)(aa * aa);
        for (uint32_t q = t0; q < npix; q += stride) {
            const uint32_t y = q / (uint32_t)a.user_w;
            const uint32_t x = q - y * (uint32_t)a.user_w;
            uint32_t out;
            if (aa == 1) {
                out = apply1<IP, COUNT>(p, a.counts[(size_t)y * (size_t)a.real_w + x], it, depth);
            } else {
                uint32_t r = 0, g = 0, b = 0;
                for (int yi = 0; yi < aa; ++yi) {
                    const uint32_t* row = a.counts + ((size_t)y * aa + yi) * (size_t)a.real_w +
                                          (size_t)x * aa;
                    for (int xi = 0; xi < aa; ++xi)
                        sample<IP, COUNT>(p, row[xi], r, g, b, it, depth);
                }
                out = pack_avg(r, g, b, aaaa);
            }
            a.rgb[q] = out;
        }
    }

    if (COUNT) {
        /* wave sum, then one atomic per block */
        for (int o = 32; o > 0; o >>= 1)
            it += __shfl_xor((unsigned long long)it, o, 64);
        const int wv = threadIdx.x >> 6;
        if ((threadIdx.x & 63) == 0)
            lsum[wv] = it;
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned long long s = 0;
            for (int w = 0; w < (int)(blockDim.x >> 6); ++w)
                s += lsum[w];
            atomicAdd(a.iters, s);
        }
    }
}

typedef void (*kfn)(colour_args);

/* the colouring pass as a launch node (mdzg_dev.h) */
extern "C" int mdzg_dev_node_colour(mdzg_gnode* node, const uint32_t* counts, int real_w,
                                    int user_w, int user_h, int aa, const uint32_t* palette,
                                    int n, int offset, double sc, int ip, uint32_t depth,
                                    uint32_t* rgb, unsigned long long* iters, int max_blocks)
{
    if (!node || !counts || !palette || !rgb || user_w < 1 || user_h < 1 || aa < 1 || n < 2 ||
        n > 256 || real_w < user_w * aa ||
        (uint64_t)user_w * (uint64_t)user_h > 0x7fffffffull || max_blocks < 1)
        return -1;
    colour_args a;
    a.counts = counts;
    a.palette = palette;
    a.rgb = rgb;
    a.iters = iters;
    a.sc = sc;
    a.mmagic = mdzg::colour::rem_magic((uint32_t)(n - 1));
    a.real_w = real_w;
    a.user_w = user_w;
    a.user_h = user_h;
    a.aa = aa;
    a.n = n;
    a.offset = offset;
    a.ip = ip;
    a.depth = depth;
    const uint32_t npix = (uint32_t)user_w * (uint32_t)user_h;
    const int aligned = ((uintptr_t)counts % 16 == 0) && ((uintptr_t)rgb % 16 == 0);
    int form = 0;
    uint32_t work = npix;
    if (aligned && aa == 1 && real_w == user_w) {
        form = 1;
        work = (npix + 3) / 4;
    } else if (aligned && aa == 2 && user_w % 2 == 0 && real_w == 2 * user_w) {
        form = 2;
        work = npix / 2;
    }
    uint32_t grid = (work + 255) / 256;
    if (grid > (uint32_t)max_blocks)
        grid = (uint32_t)max_blocks;
    if (grid < 1)
        grid = 1;
    static const kfn table[3][2][2] = {
        { { k_colour<0, false, false>, k_colour<0, false, true> },
          { k_colour<0, true, false>, k_colour<0, true, true> } },
        { { k_colour<1, false, false>, k_colour<1, false, true> },
          { k_colour<1, true, false>, k_colour<1, true, true> } },
        { { k_colour<2, false, false>, k_colour<2, false, true> },
          { k_colour<2, true, false>, k_colour<2, true, true> } },
    };
    static_assert(sizeof(colour_args) <= MDZG_GNODE_BUF, "colour_args fits a launch node");
    memset(node, 0, sizeof *node);
    node->func = (const void*)table[form][ip ? 1 : 0][iters ? 1 : 0];
    node->grid = grid;
    node->block = 256;
    node->nargs = 1;
    node->off[0] = 0;
    node->used = (int)sizeof a;
    memcpy(node->buf, &a, sizeof a);
    return 0;
}

extern "C" int mdzg_dev_node_launch(const mdzg_gnode* n, hipStream_t st);

extern "C" int mdzg_dev_launch_colour(const uint32_t* counts, int real_w, int user_w, int user_h,
                                      int aa, const uint32_t* palette, int n, int offset,
                                      double sc, int ip, uint32_t depth, uint32_t* rgb,
                                      unsigned long long* iters, int max_blocks, hipStream_t st)
{
    mdzg_gnode node;
    if (mdzg_dev_node_colour(&node, counts, real_w, user_w, user_h, aa, palette, n, offset, sc,
                             ip, depth, rgb, iters, max_blocks) != 0)
        return -1;
    return mdzg_dev_node_launch(&node, st) == 0 ? 0 : -1;
}
