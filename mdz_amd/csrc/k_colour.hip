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
 * HBM-bound: 4 aa^2 bytes read and 4 bytes written per output pixel; the
 * palette (<= 256 colours) is staged in LDS.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

struct colour_args {
    const uint32_t* counts;      /* real_w x real_h, row-major                  */
    const uint32_t* palette;     /* n colours                                   */
    uint32_t* rgb;               /* user_w x user_h                             */
    unsigned long long* iters;   /* pixel-iteration sum (added to), may be null */
    double sc;                   /* image_info.colour_scale                     */
    int real_w, user_w, user_h, aa;
    int n;                       /* pal_indexes, 2..256                         */
    int offset;                  /* pal_offset                                  */
    int ip;                      /* image_info.palette_ip                       */
    uint32_t depth;
};

__device__ __forceinline__ uint32_t u32_of(double v)
{
    return v < 9223372036854775808.0 ? (uint32_t)(uint64_t)(int64_t)v : 0u;
}

__device__ __forceinline__ uint32_t u8_of(double v)
{
    return (uint32_t)v & 0xffu;
}

/* one interpolated colour: the shared body of both interpolating paths */
__device__ __forceinline__ uint32_t blend(const uint32_t* pal, uint32_t m, double v, double cval)
{
    const double diff = cval - v;
    const double rdiff = 1.0 - diff;
    const uint32_t c1 = pal[u32_of(floor(v)) % m];
    const uint32_t c2 = pal[u32_of(cval) % m];
    const uint32_t r = u8_of(diff * (double)(c1 & 0xffu) + rdiff * (double)(c2 & 0xffu));
    const uint32_t g = u8_of(diff * (double)((c1 >> 8) & 0xffu) + rdiff * (double)((c2 >> 8) & 0xffu));
    const uint32_t b = u8_of(diff * (double)((c1 >> 16) & 0xffu) + rdiff * (double)((c2 >> 16) & 0xffu));
    return r | (g << 8) | (b << 16);
}

/* palette_apply, one sample */
__device__ __forceinline__ uint32_t colour_apply(const uint32_t* pal, uint32_t m, uint32_t raw,
                                                 double sc, int offset, int ip)
{
    if (!raw)
        return 0u;
    const double v = (double)(int32_t)raw * sc + (double)offset;
    if (ip)
        return blend(pal, m, v, ceil(v));
    return pal[u32_of(v) % m];
}

/* get_pixel_colour(raw * sc, ip) */
__device__ __forceinline__ uint32_t colour_get(const uint32_t* pal, uint32_t m, uint32_t raw,
                                               double sc, int offset, int ip)
{
    double v = (double)(int32_t)raw * sc;
    if (v != 0.0)
        v += (double)offset;
    if (ip)
        return blend(pal, m, v, ceil(v));
    if (v == 0.0)
        return 0u;
    return pal[u32_of(v) % m];
}

__global__ __launch_bounds__(256) void k_colour(colour_args a)
{
    __shared__ uint32_t lpal[256];
    __shared__ unsigned long long lsum[4];
    for (int i = threadIdx.x; i < a.n; i += blockDim.x)
        lpal[i] = a.palette[i];
    __syncthreads();

    const uint32_t npix = (uint32_t)a.user_w * (uint32_t)a.user_h;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t m = (uint32_t)(a.n - 1);
    unsigned long long it = 0;
    if (idx < npix) {
        const uint32_t y = idx / (uint32_t)a.user_w;
        const uint32_t x = idx - y * (uint32_t)a.user_w;
        const int aa = a.aa;
        uint32_t out;
        if (aa == 1) {
            const uint32_t raw = a.counts[(size_t)y * (size_t)a.real_w + x];
            it = raw ? raw : a.depth;
            out = colour_apply(lpal, m, raw, a.sc, a.offset, a.ip);
        } else {
            uint32_t r = 0, g = 0, b = 0;
            for (int yi = 0; yi < aa; ++yi) {
                const uint32_t* row = a.counts + ((size_t)y * aa + yi) * (size_t)a.real_w +
                                      (size_t)x * aa;
                for (int xi = 0; xi < aa; ++xi) {
                    const uint32_t raw = row[xi];
                    it += raw ? raw : a.depth;
                    const uint32_t c = colour_get(lpal, m, raw, a.sc, a.offset, a.ip);
                    r += c & 0xffu;
                    g += (c >> 8) & 0xffu;
                    b += (c >> 16) & 0xffu;
                }
            }
            const uint32_t aaaa = (uint32_t)(aa * aa);
            out = (r / aaaa) | ((g / aaaa) << 8) | ((b / aaaa) << 16);
        }
        a.rgb[idx] = out;
    }
    if (a.iters) {
        /* wave sum, then one atomic per block */
        for (int o = 32; o > 0; o >>= 1)
            it += __shfl_xor(it, o, 64);
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

extern "C" int mdzg_dev_launch_colour(const uint32_t* counts, int real_w, int user_w, int user_h,
                                      int aa, const uint32_t* palette, int n, int offset,
                                      double sc, int ip, uint32_t depth, uint32_t* rgb,
                                      unsigned long long* iters, hipStream_t st)
{
    if (!counts || !palette || !rgb || user_w < 1 || user_h < 1 || aa < 1 || n < 2 || n > 256 ||
        real_w < user_w * aa || (uint64_t)user_w * (uint64_t)user_h > 0x7fffffffull)
        return -1;
    colour_args a;
    a.counts = counts;
    a.palette = palette;
    a.rgb = rgb;
    a.iters = iters;
    a.sc = sc;
    a.real_w = real_w;
    a.user_w = user_w;
    a.user_h = user_h;
    a.aa = aa;
    a.n = n;
    a.offset = offset;
    a.ip = ip;
    a.depth = depth;
    const uint32_t npix = (uint32_t)user_w * (uint32_t)user_h;
    hipLaunchKernelGGL(k_colour, dim3((npix + 255) / 256), dim3(256), 0, st, a);
    return hipGetLastError() == hipSuccess ? 0 : -1;
}
