/*
 * colour_math.h -- the per-sample colour arithmetic of k_colour.hip, for the
 * device and (for checks) the host: MDZ's palette_apply one-sample body
 * (src/palette.c:417-440 interpolating, :447-458 plain) and get_pixel_colour
 * (src/palette.c:466-502).
 *
 * Every step is MDZ's in IEEE double with no fused operations (build with
 * -ffp-contract=off); the double-to-integer conversions are the ones x86-64
 * performs, where MDZ's build runs:
 *   (guint32)v  the low 32 bits of the 64-bit truncation, 0 where that
 *               truncation overflows (|v| >= 2^63, infinities);
 *   (guint8)v   the low byte of the 32-bit truncation (v is in [0, 256) here).
 * palette_apply's interpolating path takes (int)floor(v) and (int)ceil(v),
 * equal to the (guint32) forms below 2^31; above, MDZ's palette index is
 * negative (undefined) and these functions give get_pixel_colour's colour.
 */
#ifndef MDZG_COLOUR_MATH_H
#define MDZG_COLOUR_MATH_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define MC_HD __host__ __device__ __forceinline__
#else
#define MC_HD static inline
#endif

namespace mdzg {
namespace colour {

/* (guint32)v */
MC_HD uint32_t u32_of(double v)
{
    if (v >= 0.0 && v < 4294967296.0)
        return (uint32_t)v;
    if (v > -9223372036854775808.0 && v < 9223372036854775808.0)
        return (uint32_t)(uint64_t)(int64_t)v;
    return 0u;
}

/* (guint8)v for v in [0, 256) */
MC_HD uint32_t u8_of(double v)
{
    return (uint32_t)v & 0xffu;
}

/* Remainder by d (1 <= d < 2^32) with one 64-bit constant
   M = floor((2^64 - 1) / d) + 1 (mod 2^64): a % d = high 64 bits of
   (M a mod 2^64) * d, for every 32-bit a (Lemire, Kaser, Kurz 2019). */
MC_HD uint64_t rem_magic(uint32_t d)
{
    return ~(uint64_t)0 / d + 1;
}

MC_HD uint32_t mulhi32(uint32_t a, uint32_t b)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __umulhi(a, b);
#else
    return (uint32_t)(((uint64_t)a * b) >> 32);
#endif
}

MC_HD uint32_t rem(uint32_t a, uint64_t M, uint32_t d)
{
    const uint64_t low = M * (uint64_t)a;
    const uint64_t t = (uint64_t)(uint32_t)(low >> 32) * d + mulhi32((uint32_t)low, d);
    return (uint32_t)(t >> 32);
}

/* A palette as the colouring reads it.  (Channels are converted to double
   per use, as MDZ does; a table of converted channels in LDS measured slower
   on the device.) */
struct palette_view {
    const uint32_t* pal;
    uint64_t M;          /* rem_magic(m) */
    uint32_t m;          /* pal_indexes - 1 */
};

/* the interpolated colour: weights ceil(v) - v on the entry at floor(v) and
   1 - that on the entry at ceil(v), per channel truncated to 8 bits */
MC_HD uint32_t blend(const palette_view& P, double v)
{
    const double cval = ceil(v);
    const double diff = cval - v;
    const double rdiff = 1.0 - diff;
    const uint32_t c1 = P.pal[rem(u32_of(floor(v)), P.M, P.m)];
    const uint32_t c2 = P.pal[rem(u32_of(cval), P.M, P.m)];
    const uint32_t r = u8_of(diff * (double)(c1 & 0xffu) + rdiff * (double)(c2 & 0xffu));
    const uint32_t g =
        u8_of(diff * (double)((c1 >> 8) & 0xffu) + rdiff * (double)((c2 >> 8) & 0xffu));
    const uint32_t b =
        u8_of(diff * (double)((c1 >> 16) & 0xffu) + rdiff * (double)((c2 >> 16) & 0xffu));
    return r | (g << 8) | (b << 16);
}

/* palette_apply for one count (raw_data is int); a zero count is black.  The
   colour of the count is formed either way and selected, so a wave of mixed
   zero and non-zero counts takes no branch. */
template <bool IP>
MC_HD uint32_t colour_apply(const palette_view& P, uint32_t raw, double sc, int offset)
{
    const double v = (double)(int32_t)raw * sc + (double)offset;
    const uint32_t c = IP ? blend(P, v) : P.pal[rem(u32_of(v), P.M, P.m)];
    return raw ? c : 0u;
}

/* get_pixel_colour(raw * sc, ip): the offset is added only to a non-zero
   value; a zero value is black without interpolation and the blend of
   entry 0 with itself (that is, entry 0) with it */
template <bool IP>
MC_HD uint32_t colour_get(const palette_view& P, uint32_t raw, double sc, int offset)
{
    const double v0 = (double)(int32_t)raw * sc;
    const double v = v0 != 0.0 ? v0 + (double)offset : v0;
    if (IP)
        return blend(P, v);
    const uint32_t c = P.pal[rem(u32_of(v), P.M, P.m)];
    return v != 0.0 ? c : 0u;
}

}  // namespace colour
}  // namespace mdzg

#endif
