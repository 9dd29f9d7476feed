/*
 * mp_rn.h -- multi-word binary floating point with round-to-nearest-even at a
 * run-time precision p, for the escape-time kernels.
 *
 * Every add, sub, mul and sqr returns exactly what MPFR returns for
 * mpfr_add / mpfr_sub / mpfr_mul with MPFR_RNDN at precision p (the operation
 * sequence of src/frac_mandel.c:33-51 and its siblings).  That is: the exact
 * result, rounded once to p significant bits, ties to even.  No exponent
 * range limit applies inside the escape loop (|values| < 2^8; the smallest
 * exponents stay far above the int32 floor), so MPFR's own range never bites.
 *
 * Representation (struct mp<N>):
 *   m[0..N-1]  32-bit words of the significand M, least significant first;
 *              nonzero values are normalised, 2^(32N-1) <= M < 2^(32N)
 *   e          exponent, value = (-1)^s * M * 2^(e - 32N)   (MPFR convention:
 *              significand in [1/2, 1) times 2^e)
 *   s          sign bit
 * Zero is M == 0 with e = EZERO (far below any reachable exponent), so the
 * alignment in add() treats it as vanishing.  The sign of a zero never
 * reaches a count: every consumer of a zero either adds a nonzero c to it or
 * squares it.
 *
 * p <= 32N.  After rounding, the low 32N - p bits of M are zero.  kclr = 32N - p
 * is uniform across a wave (one tile, one precision): kclr == 0 takes the
 * word-aligned rounding path, kclr > 0 the masked one.
 *
 * gfx950 mapping: a 32x32+64 multiply-accumulate is one v_mad_u64_u32 whose
 * carry-out feeds one v_addc_co_u32 into the third accumulator word, issued
 * in groups that keep the carry's read two instructions after its write
 * (macg below; the host form computes the same function in portable C).  Word carry chains use v_add_co/v_addc_co via __builtin_addc.
 * Rare cases (carry through a rounding increment, cancellation of a whole
 * word, an alignment shift of whole words) run under wave-wide votes, so a wave
 * only pays for them when one of its lanes needs them.  Built for the host
 * (tests only) the vote is the lane's own predicate.
 */
#ifndef MDZG_MP_RN_H
#define MDZG_MP_RN_H

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define MP_HD __host__ __device__ __forceinline__
#else
#define MP_HD static inline
#endif

namespace mdzg {

static constexpr int32_t EZERO = -(1 << 30);

template <int N>
struct mp {
    uint32_t m[N];
    int32_t e;
    uint32_t s;
};

MP_HD bool wave_any(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* the vote's lane mask is the compare's own scalar result: no VALU work */
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
}

/* A lane mask copied by the scalar unit.  A compare writes its lane mask
   with a VALU instruction, and gfx950 needs two wait states before a VALU
   instruction reads an SGPR that a VALU instruction wrote; the compiler
   inserts them for its own code but not before an asm block.  The asm
   blocks below that read a mask read this copy, which an SALU instruction
   wrote: that read needs no wait. */
MP_HD uint64_t scalar_copy(uint64_t m)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint64_t r;
    asm("s_mov_b64 %0, %1" : "=s"(r) : "s"(m));
    return r;
#else
    return m;
#endif
}

/* lane mask of a per-lane predicate (device, a scalar copy for the asm
   readers) / the predicate itself (host) */
MP_HD uint64_t lane_mask(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return scalar_copy(__builtin_amdgcn_ballot_w64(c));
#else
    return c ? 1u : 0u;
#endif
}

/* m ? y : x word by word, opaque to the optimiser on the device so that the
   swap of two operands stays in registers (a plain select of two arrays is
   turned into a select of pointers and a copy through private memory) */
MP_HD uint32_t sel_mask(uint64_t m, uint32_t x, uint32_t y)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t r;
    asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(x), "v"(y), "s"(m));
    return r;
#else
    return m ? y : x;
#endif
}

/* (hi:lo) >> s, low 32 bits, for 0 <= s < 32 (v_alignbit_b32) */
MP_HD uint32_t funnel_r(uint32_t hi, uint32_t lo, uint32_t s)
{
#if defined(__HIP_DEVICE_COMPILE__)
    /* the builtin keeps word arrays as words (a 64-bit shift of two adjacent
       words is fused into a 64-bit load that pins the array in memory) */
    return __builtin_amdgcn_alignbit(hi, lo, s & 31u);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
#endif
}

/* a + b + cin, carry to cout (v_add_co_u32 / v_addc_co_u32) */
MP_HD uint32_t add_cc(uint32_t a, uint32_t b, uint32_t cin, uint32_t& cout)
{
#if defined(__clang__)
    unsigned int co;
    const uint32_t r = __builtin_addc(a, b, cin, &co);
    cout = co;
    return r;
#else
    const uint64_t t = (uint64_t)a + b + cin;
    cout = (uint32_t)(t >> 32);
    return (uint32_t)t;
#endif
}

/* a - b - bin, borrow to bout (v_sub_co_u32 / v_subb_co_u32) */
MP_HD uint32_t sub_bb(uint32_t a, uint32_t b, uint32_t bin, uint32_t& bout)
{
#if defined(__clang__)
    unsigned int bo;
    const uint32_t r = __builtin_subc(a, b, bin, &bo);
    bout = bo;
    return r;
#else
    const uint64_t t = (uint64_t)a - b - bin;
    bout = (uint32_t)(t >> 63);
    return (uint32_t)t;
#endif
}

/* number of power-of-two stages of a shifter covering shifts 0..n */
constexpr int stages(int n)
{
    return n <= 0 ? 0 : 1 + stages(n >> 1);
}

MP_HD uint32_t clz32(uint32_t x)
{
    return x ? (uint32_t)__builtin_clz(x) : 32u;
}

template <int N>
MP_HD void set_zero(mp<N>& r)
{
#pragma unroll
    for (int j = 0; j < N; ++j) r.m[j] = 0;
    r.e = EZERO;
    r.s = 0;
}

template <int N>
MP_HD bool is_zero(const mp<N>& a)
{
    return a.m[N - 1] == 0;
}

/* ------------------------------------------------------------------ */
/* Rounding of a normalised candidate R (N words) with a guard word G   */
/* (the 32 bits right below R) and a sticky flag S (anything below G).  */
/* ------------------------------------------------------------------ */

/* R += inc << bit at word w0 (runtime uniform), carries propagate upward;
   a carry out of the top word means R was all ones: R = 2^(32N-1), e + 1. */
template <int N>
MP_HD void add_ulp(uint32_t* R, int32_t& e, uint32_t inc_word0)
{
    uint32_t c;
    R[0] = add_cc(R[0], inc_word0, 0u, c);
    if (__builtin_expect(wave_any(c != 0), 0)) {      /* R[0] wrapped: rare */
#pragma unroll
        for (int j = 1; j < N; ++j) R[j] = add_cc(R[j], 0u, c, c);
        if (c) {
            R[N - 1] = 0x80000000u;
            e += 1;
        }
    }
}

/* R += 1 on the lanes where inc holds, carries ripple upward as in add_ulp.
   On the device the compare's lane mask is the carry-in of one
   v_addc_co_u32, and its carry-out mask decides the rare ripple. */
template <int N>
MP_HD void add_one_if(uint32_t* R, int32_t& e, bool inc)
{
#if defined(__HIP_DEVICE_COMPILE__)
    const uint64_t m = scalar_copy(__builtin_amdgcn_ballot_w64(inc));
    uint64_t co;
    uint32_t r0;
    asm("v_addc_co_u32 %0, %1, 0, %2, %3" : "=v"(r0), "=s"(co) : "v"(R[0]), "s"(m));
    const uint32_t old = R[0];
    R[0] = r0;
    if (__builtin_expect(co != 0, 0)) {       /* R[0] wrapped on some lane: rare */
        uint32_t c = (uint32_t)(r0 < old);
#pragma unroll
        for (int j = 1; j < N; ++j) R[j] = add_cc(R[j], 0u, c, c);
        if (c) {
            R[N - 1] = 0x80000000u;
            e += 1;
        }
    }
#else
    add_ulp<N>(R, e, inc ? 1u : 0u);
#endif
}

/* S is 0 or 1.  Nearest-even rounds up when G exceeds the half 2^31, or
   equals it with a sticky bit or an odd last bit: with x = S | lsb (0 or 1)
   that is G > 2^31 - x. */
template <int N>
MP_HD void round_full(uint32_t* R, uint32_t G, uint32_t S, int32_t& e)
{
    const uint32_t x = (R[0] & 1u) | S;
    add_one_if<N>(R, e, G > 0x80000000u - x);
}

/* true when the predicate holds on every lane of the wave (device) / on
   this value (host) */
MP_HD bool wave_all(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(!c) == 0;
#else
    return c;
#endif
}

/* kclr = 32N - p in [1, 32N-1]: clear the low kclr bits of R, round at bit kclr */
template <int N>
MP_HD void round_masked(uint32_t* R, uint32_t G, uint32_t S, int32_t& e, int kclr)
{
    if (N == 1 || wave_all(kclr < 32)) {
        /* the rounding position lies in the lowest word (p = 80, 100, 176,
           184, ...; always for N = 1, where p >= 2): only word 0 is masked,
           the increment ripples upward as in add_ulp */
        const uint32_t rm = 1u << (kclr - 1);            /* round bit */
        const uint32_t rbit = R[0] & rm;
        const uint32_t st = G | S | (R[0] & (rm - 1u));  /* sticky    */
        const uint32_t lsb = R[0] & (rm << 1);           /* kept lsb  */
        R[0] &= ~((rm << 1) - 1u);
        const uint32_t inc = (uint32_t)(rbit != 0) & (uint32_t)((st != 0) | (lsb != 0));
        add_ulp<N>(R, e, inc << kclr);
        return;
    }
    const int wi = kclr >> 5, bi = kclr & 31;            /* lsb of the kept part */
    const int rr = kclr - 1, rw = rr >> 5, rb = rr & 31; /* round bit            */
    uint32_t rbit = 0, st = G | S, lsb = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t keep = (j > wi) ? 0xffffffffu : (j == wi ? (0xffffffffu << bi) : 0u);
        const uint32_t rm = (j == rw) ? (1u << rb) : 0u;
        const uint32_t sm = (j < rw) ? 0xffffffffu : (j == rw ? ((1u << rb) - 1u) : 0u);
        const uint32_t lm = (j == wi) ? (1u << bi) : 0u;
        rbit |= R[j] & rm;
        st |= R[j] & sm;
        lsb |= R[j] & lm;
        R[j] &= keep;
    }
    const uint32_t inc = (uint32_t)(rbit != 0) & (uint32_t)((st != 0) | (lsb != 0));
    /* add inc << kclr: word wi gets inc << bi, carries ripple upward
       (words below wi are zero and receive nothing) */
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        const uint32_t add = (j == wi) ? (inc << bi) : 0u;
        const uint32_t cin = (j > wi) ? c : 0u;
        R[j] = add_cc(R[j], add, cin, c);
    }
    if (c) {
        R[N - 1] = 0x80000000u;
        e += 1;
    }
}

template <int N>
MP_HD void round_p(uint32_t* R, uint32_t G, uint32_t S, int32_t& e, int kclr)
{
    if (kclr == 0)
        round_full<N>(R, G, S, e);
    else
        round_masked<N>(R, G, S, e, kclr);
}

/* ------------------------------------------------------------------ */
/* Significand product, column by column (product scanning).            */
/* Outputs the top N+1 words of the 2N-word product (H[0] = word N-1,   */
/* the guard word; H[N] = word 2N-1) and the OR of words 0..N-2.        */
/* Each column sum is kept in a 96-bit accumulator (c2 : acc).          */
/* ------------------------------------------------------------------ */

/* Multiply-accumulate groups.  A product a*b joins the column sum with one
   v_mad_u64_u32 (acc += a*b, its carry-out to an SGPR pair) and one
   v_addc_co_u32 that adds that carry into the third word c2.  gfx950 needs
   two wait states between a VALU write of an SGPR and a VALU read of it,
   and the compiler inserts none inside or before an asm block.  A group of
   G >= 3 products therefore issues its G multiply-adds, each with its own
   SGPR pair, before its G carry adds: every carry is read at least two
   instructions after it was written.  Groups of 1 and 2 fill the gap with
   s_nop.  FIRST starts c2 from zero (a column's first group). */
template <int G, bool FIRST>
MP_HD void macg(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
#if defined(__HIP_DEVICE_COMPILE__)
    ;
template <>
MP_HD void macg<1, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0;
    asm("v_mad_u64_u32 %0, %2, %3, %4, %0\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2"
        : "+v"(acc), "=v"(c2), "=&s"(s0)
        : "v"(a[0]), "v"(b[0]));
}
template <>
MP_HD void macg<1, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0;
    asm("v_mad_u64_u32 %0, %2, %3, %4, %0\n\t"
        "s_nop 1\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2"
        : "+v"(acc), "+v"(c2), "=&s"(s0)
        : "v"(a[0]), "v"(b[0]));
}
template <>
MP_HD void macg<2, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\t"
        "v_mad_u64_u32 %0, %3, %6, %7, %0\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3"
        : "+v"(acc), "=v"(c2), "=&s"(s0), "=&s"(s1)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]));
}
template <>
MP_HD void macg<2, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1;
    asm("v_mad_u64_u32 %0, %2, %4, %5, %0\n\t"
        "v_mad_u64_u32 %0, %3, %6, %7, %0\n\t"
        "s_nop 0\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3"
        : "+v"(acc), "+v"(c2), "=&s"(s0), "=&s"(s1)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]));
}
template <>
MP_HD void macg<3, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2;
    asm("v_mad_u64_u32 %0, %2, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, %3, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, %4, %9, %10, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4"
        : "+v"(acc), "=v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]));
}
template <>
MP_HD void macg<3, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2;
    asm("v_mad_u64_u32 %0, %2, %5, %6, %0\n\t"
        "v_mad_u64_u32 %0, %3, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, %4, %9, %10, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4"
        : "+v"(acc), "+v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]));
}
template <>
MP_HD void macg<4, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3;
    asm("v_mad_u64_u32 %0, %2, %6, %7, %0\n\t"
        "v_mad_u64_u32 %0, %3, %8, %9, %0\n\t"
        "v_mad_u64_u32 %0, %4, %10, %11, %0\n\t"
        "v_mad_u64_u32 %0, %5, %12, %13, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5"
        : "+v"(acc), "=v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]));
}
template <>
MP_HD void macg<4, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3;
    asm("v_mad_u64_u32 %0, %2, %6, %7, %0\n\t"
        "v_mad_u64_u32 %0, %3, %8, %9, %0\n\t"
        "v_mad_u64_u32 %0, %4, %10, %11, %0\n\t"
        "v_mad_u64_u32 %0, %5, %12, %13, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5"
        : "+v"(acc), "+v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]));
}
template <>
MP_HD void macg<5, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3, s4;
    asm("v_mad_u64_u32 %0, %2, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, %3, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, %4, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, %5, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, %6, %15, %16, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5\n\t"
        "v_addc_co_u32 %1, %6, 0, %1, %6"
        : "+v"(acc), "=v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3), "=&s"(s4)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]), "v"(a[4]), "v"(b[4]));
}
template <>
MP_HD void macg<5, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3, s4;
    asm("v_mad_u64_u32 %0, %2, %7, %8, %0\n\t"
        "v_mad_u64_u32 %0, %3, %9, %10, %0\n\t"
        "v_mad_u64_u32 %0, %4, %11, %12, %0\n\t"
        "v_mad_u64_u32 %0, %5, %13, %14, %0\n\t"
        "v_mad_u64_u32 %0, %6, %15, %16, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5\n\t"
        "v_addc_co_u32 %1, %6, 0, %1, %6"
        : "+v"(acc), "+v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3), "=&s"(s4)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]), "v"(a[4]), "v"(b[4]));
}
template <>
MP_HD void macg<6, true>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3, s4, s5;
    asm("v_mad_u64_u32 %0, %2, %8, %9, %0\n\t"
        "v_mad_u64_u32 %0, %3, %10, %11, %0\n\t"
        "v_mad_u64_u32 %0, %4, %12, %13, %0\n\t"
        "v_mad_u64_u32 %0, %5, %14, %15, %0\n\t"
        "v_mad_u64_u32 %0, %6, %16, %17, %0\n\t"
        "v_mad_u64_u32 %0, %7, %18, %19, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, 0, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5\n\t"
        "v_addc_co_u32 %1, %6, 0, %1, %6\n\t"
        "v_addc_co_u32 %1, %7, 0, %1, %7"
        : "+v"(acc), "=v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3), "=&s"(s4), "=&s"(s5)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]), "v"(a[4]), "v"(b[4]), "v"(a[5]), "v"(b[5]));
}
template <>
MP_HD void macg<6, false>(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    uint64_t s0, s1, s2, s3, s4, s5;
    asm("v_mad_u64_u32 %0, %2, %8, %9, %0\n\t"
        "v_mad_u64_u32 %0, %3, %10, %11, %0\n\t"
        "v_mad_u64_u32 %0, %4, %12, %13, %0\n\t"
        "v_mad_u64_u32 %0, %5, %14, %15, %0\n\t"
        "v_mad_u64_u32 %0, %6, %16, %17, %0\n\t"
        "v_mad_u64_u32 %0, %7, %18, %19, %0\n\t"
        "v_addc_co_u32 %1, %2, 0, %1, %2\n\t"
        "v_addc_co_u32 %1, %3, 0, %1, %3\n\t"
        "v_addc_co_u32 %1, %4, 0, %1, %4\n\t"
        "v_addc_co_u32 %1, %5, 0, %1, %5\n\t"
        "v_addc_co_u32 %1, %6, 0, %1, %6\n\t"
        "v_addc_co_u32 %1, %7, 0, %1, %7"
        : "+v"(acc), "+v"(c2), "=&s"(s0), "=&s"(s1), "=&s"(s2), "=&s"(s3), "=&s"(s4), "=&s"(s5)
        : "v"(a[0]), "v"(b[0]), "v"(a[1]), "v"(b[1]), "v"(a[2]), "v"(b[2]), "v"(a[3]), "v"(b[3]), "v"(a[4]), "v"(b[4]), "v"(a[5]), "v"(b[5]));
}

#else
{
    if (FIRST)
        c2 = 0;
    for (int k = 0; k < G; ++k) {
        const uint64_t p = (uint64_t)a[k] * b[k];
        const uint64_t t = acc + p;
        c2 += (uint32_t)(t < p);
        acc = t;
    }
}
#endif

/* (c2 : acc) += sum of a[t] * b[t] over t < n, in groups of at most 6
   (4 while more than 6 are left, so no group but a column's only one is
   smaller than 3); FIRST: c2 starts from zero */
template <int n, bool FIRST>
MP_HD void mac_col(const uint32_t* a, const uint32_t* b, uint64_t& acc, uint32_t& c2)
{
    constexpr int g = n > 6 ? 4 : n;
    macg<g, FIRST>(a, b, acc, c2);
    if constexpr (n > g)
        mac_col<n - g, false>(a + g, b + g, acc, c2);
}

/* one product added to a running column (the squares' diagonal, the GMP
   path's loops) */
MP_HD void mac96(uint32_t a, uint32_t b, uint64_t& acc, uint32_t& c2)
{
    macg<1, false>(&a, &b, acc, c2);
}

/* the first product of a column: c2 starts from its carry */
MP_HD void mac96_first(uint32_t a, uint32_t b, uint64_t& acc, uint32_t& c2)
{
    macg<1, true>(&a, &b, acc, c2);
}

/* columns k = K .. 2N-2 of a * b: products a[i] b[k-i] over 0 <= i, k-i < N */
template <int N, int K>
MP_HD void mul_cols(const uint32_t* a, const uint32_t* b, uint32_t* H, uint64_t& acc,
                    uint32_t& st)
{
    if constexpr (K < 2 * N - 1) {
        constexpr int i0 = K < N ? 0 : K - N + 1;
        constexpr int n = K < N ? K + 1 : 2 * N - 1 - K;
        uint32_t ca[n], cb[n];
#pragma unroll
        for (int t = 0; t < n; ++t) {
            ca[t] = a[i0 + t];
            cb[t] = b[K - i0 - t];
        }
        uint32_t c2 = 0;
        if constexpr (K == 2 * N - 2)
            acc += (uint64_t)ca[0] * cb[0];   /* top column: the sum fits 64 bits */
        else
            mac_col<n, true>(ca, cb, acc, c2);
        if (K < N - 1)
            st |= (uint32_t)acc;
        else
            H[K - (N - 1)] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)c2 << 32);
        mul_cols<N, K + 1>(a, b, H, acc, st);
    }
}

/* K0 > 0 leaves out the columns below K0 (a short product, see mul_rn);
   S is then meaningless */
template <int N, int K0>
MP_HD void mant_mul(const uint32_t* a, const uint32_t* b, uint32_t* H, uint32_t& S)
{
    uint64_t acc = 0;
    uint32_t st = 0;
    mul_cols<N, K0>(a, b, H, acc, st);
    H[N] = (uint32_t)acc;
    S = (uint32_t)(st != 0);
}

/* Words k = K .. 2N-1 of a * a with each off-diagonal product formed once.
   Column k of the off-diagonal sum O = sum over i < j of a[i] a[j]
   2^(32(i+j)) is product-scanned as in mul_cols; as soon as its word o is
   final, word k of the square, 2 O + sum of a[i]^2 2^(64 i), follows from o,
   the previous word of O (the doubling shifts its top bit in), one word of
   a diagonal square and the carry c of the pass: two instructions a word
   where doubling each column's 96-bit sum took six.  The columns below the
   first K are left out of both sums, as in mul_cols (a diagonal square whose
   column 2i is left out is left out whole). */
template <int N, int K>
MP_HD void sqr_cols(const uint32_t* a, uint32_t* H, uint64_t& acc, uint32_t& prev,
                    uint32_t& c, uint32_t& dhi, uint32_t& st)
{
    if constexpr (K < 2 * N) {
        uint32_t o;
        if constexpr (K <= 2 * N - 3) {
            constexpr int i0 = K < N ? 0 : K - N + 1;
            constexpr int i1 = (K + 1) / 2 - 1;     /* largest i with i < K - i */
            constexpr int n = i1 >= i0 ? i1 - i0 + 1 : 0;
            uint32_t c2 = 0;
            if constexpr (n > 0) {
                uint32_t ca[n], cb[n];
#pragma unroll
                for (int t = 0; t < n; ++t) {
                    ca[t] = a[i0 + t];
                    cb[t] = a[K - i0 - t];
                }
                mac_col<n, true>(ca, cb, acc, c2);
            }
            o = (uint32_t)acc;
            acc = (acc >> 32) | ((uint64_t)c2 << 32);
        } else {
            /* no off-diagonal products: O's last two words (O < 2^(64N-1)) */
            o = (K == 2 * N - 2) ? (uint32_t)acc : (uint32_t)(acc >> 32);
        }
        uint32_t d;
        if constexpr ((K & 1) == 0) {
            const uint64_t sq = (uint64_t)a[K >> 1] * a[K >> 1];
            d = (uint32_t)sq;
            dhi = (uint32_t)(sq >> 32);
        } else {
            d = dhi;
        }
        const uint32_t w = add_cc(funnel_r(o, prev, 31u), d, c, c);
        if (K < N - 1)
            st |= w;
        else
            H[K - (N - 1)] = w;
        prev = o;
        sqr_cols<N, K + 1>(a, H, acc, prev, c, dhi, st);
    }
}

/* Square with the off-diagonal products formed once (n(n-1)/2
   multiply-accumulates plus n squares). */
template <int N, int K0>
MP_HD void mant_sqr_sym(const uint32_t* a, uint32_t* H, uint32_t& S)
{
    uint64_t acc = 0;
    uint32_t prev = 0, c = 0, dhi = 0, st = 0;
    sqr_cols<N, K0>(a, H, acc, prev, c, dhi, st);
    S = (uint32_t)(st != 0);
}

/* The exact square (all columns), for the rare fallback of the short one:
   each column's off-diagonal sum doubled into the running column sum.  The
   same result as the doubling pass above, which needs more registers when
   it runs over all 2N columns: with it here the ragged 16- and 24-word
   persistent kernels spilled 260 and 316 bytes a lane (this form: 0 and
   36). */
template <int N, int K>
MP_HD void sqr_cols_exact(const uint32_t* a, uint32_t* H, uint64_t& acc, uint32_t& st)
{
    if constexpr (K < 2 * N - 1) {
        constexpr int i0 = K < N ? 0 : K - N + 1;
        constexpr int i1 = (K + 1) / 2 - 1;         /* largest i with i < K - i */
        constexpr int n = i1 >= i0 ? i1 - i0 + 1 : 0;
        uint32_t c2 = 0;
        if constexpr (n > 0) {
            uint32_t ca[n], cb[n];
#pragma unroll
            for (int t = 0; t < n; ++t) {
                ca[t] = a[i0 + t];
                cb[t] = a[K - i0 - t];
            }
            uint64_t o = 0;
            uint32_t o2;
            mac_col<n, true>(ca, cb, o, o2);
            /* acc += 2 * (o2:o) */
            const uint32_t d0 = (uint32_t)o << 1;
            const uint32_t d1 = funnel_r((uint32_t)(o >> 32), (uint32_t)o, 31u);
            const uint32_t d2 = funnel_r(o2, (uint32_t)(o >> 32), 31u);
            uint32_t cc;
            const uint32_t l0 = add_cc((uint32_t)acc, d0, 0u, cc);
            const uint32_t l1 = add_cc((uint32_t)(acc >> 32), d1, cc, cc);
            c2 = d2 + cc;
            acc = ((uint64_t)l1 << 32) | l0;
        }
        if constexpr (K == 2 * N - 2)
            acc += (uint64_t)a[K >> 1] * a[K >> 1];   /* top column: fits 64 bits */
        else if constexpr ((K & 1) == 0)
            mac96(a[K >> 1], a[K >> 1], acc, c2);
        if (K < N - 1)
            st |= (uint32_t)acc;
        else
            H[K - (N - 1)] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)c2 << 32);
        sqr_cols_exact<N, K + 1>(a, H, acc, st);
    }
}

template <int N>
MP_HD void mant_sqr_exact(const uint32_t* a, uint32_t* H, uint32_t& S)
{
    uint64_t acc = 0;
    uint32_t st = 0;
    sqr_cols_exact<N, 0>(a, H, acc, st);
    H[N] = (uint32_t)acc;
    S = (uint32_t)(st != 0);
}

/* Short products.  From MDZG_SHORT_MIN words up, mul_rn first forms only
   the columns k >= K0 = N - 2 of the 2N-word product.  The columns left out
   sum to D < K0 * 2^(32(K0+1)) (each column k holds k+1 products below
   2^64), i.e. less than N - 2 units of the guard word H[0], so the exact
   product lies in [T, T + (N-1) units of H[0]) where T is the short
   product with its discarded column-K0 word; after the one-bit
   normalisation, within 2(N-1) units.  Round-to-nearest of every value in
   that interval is the same unless the interval reaches a midpoint:
   for p = 32N the midpoint is H[0] = 2^31, for ragged p it lies on a word
   boundary of H[0] (H[0] = 0 with the kept part's low bits 10..0, or a
   carry out of H[0] into low bits 01..1).  Those lanes (a few in 2^32) take
   the full product; the others round T with a nonzero sticky bit, which is
   exact because the exact value is not a midpoint.  MPFR's own mulhigh
   path rests on the same bound-and-check argument. */
#ifndef MDZG_SHORT_MIN
#define MDZG_SHORT_MIN 3
#endif
#if defined(MDZG_COUNT_SHORT_FALLBACK) && !defined(__HIP_DEVICE_COMPILE__)
extern long mdzg_short_fallbacks;   /* host tests: how often the check failed */
#define MDZG_NOTE_SHORT_FALLBACK() (++mdzg_short_fallbacks)
#else
#define MDZG_NOTE_SHORT_FALLBACK() ((void)0)
#endif
template <int N>
constexpr int short_col()
{
    return N >= MDZG_SHORT_MIN ? N - 2 : 0;
}

template <int N, bool SQR, int K0>
MP_HD void mant_product(const uint32_t* a, const uint32_t* b, uint32_t* H, uint32_t& S)
{
    /* short squares by the doubling pass from 6 words (+4-10 % on 6- and
       8-word frames; neutral at 4), exact ones by columns from 12 */
    if (SQR && N >= 12 && K0 == 0)
        mant_sqr_exact<N>(a, H, S);
    else if (SQR && N >= 6 && K0 > 0)
        mant_sqr_sym<N, K0>(a, H, S);
    else
        mant_mul<N, K0>(a, SQR ? a : b, H, S);
}

/* product of normalised values is >= 2^(64N-2): at most one left shift,
   skipped by a wave none of whose lanes needs it */
template <int N>
MP_HD void norm_product(uint32_t* H, int32_t& e)
{
    const uint32_t dec = (H[N] >> 31) ^ 1u;       /* 1 where a shift is needed */
    if (wave_any(dec != 0)) {
#pragma unroll
        for (int j = N; j >= 1; --j)
            H[j] = dec ? funnel_r(H[j], H[j - 1], 31u) : H[j];
        H[0] = dec ? (H[0] << 1) : H[0];
        e -= (int32_t)dec;
    }
}

/* r = RN_p(a * b) (SQR: b ignored, a*a) */
template <int N, bool SQR>
MP_HD void mul_rn(mp<N>& r, const mp<N>& a, const mp<N>& b, int kclr)
{
    constexpr int K0 = short_col<N>();
    uint32_t H[N + 1];
    uint32_t S;
    mant_product<N, SQR, K0>(a.m, b.m, H, S);
    const int32_t e0 = SQR ? (a.e + a.e) : (a.e + b.e);
    int32_t e = e0;
    const uint32_t s = SQR ? 0u : (a.s ^ b.s);
    norm_product<N>(H, e);
    if (K0 > 0) {
        constexpr uint32_t U = 2u * N;
        const bool unsure = (kclr == 0) ? ((uint32_t)(0x80000000u - H[0]) < U)
                                        : ((uint32_t)(H[0] + (U - 1u)) < U);
        S = 1;
        /* marked unlikely, the exact product is laid out past the kernel's
           loops, which keeps the loop's instructions contiguous.  Placed
           between them (the compiler's choice without the mark), the
           exact products left the 1024-bit loop with 1.2e8
           instruction-cache misses on a 737,280-pixel band against 6e4
           out of line, and 15 % slower. */
        if (__builtin_expect(wave_any(unsure), 0)) {
            MDZG_NOTE_SHORT_FALLBACK();
            mant_product<N, SQR, 0>(a.m, b.m, H, S);
            e = e0;
            norm_product<N>(H, e);
        }
    }
    round_p<N>(H + 1, H[0], S, e, kclr);
    const bool z = (H[N] == 0);
#pragma unroll
    for (int j = 0; j < N; ++j) r.m[j] = H[j + 1];
    r.e = z ? EZERO : e;
    r.s = z ? 0u : s;
}

/* ------------------------------------------------------------------ */
/* r = RN_p(a + (-1)^flipb * b)                                         */
/* KIND ADD_ANY: any signs.  KIND SUB_NONNEG: a - b for a, b >= 0 (the  */
/* squares' difference of the escape loop): a subtraction of magnitudes */
/* whatever the swap, so the carry chain is a plain borrow chain and    */
/* the same-sign carry step cannot occur.  flipb is ignored there.      */
/* ------------------------------------------------------------------ */
enum { ADD_ANY = 0, SUB_NONNEG = 1 };

template <int N, int KIND = ADD_ANY, bool WORD_ALIGNED = false>
MP_HD void add_rn(mp<N>& r, const mp<N>& a_in, const mp<N>& b_in, uint32_t flipb, int kclr)
{
    constexpr bool SUBNN = KIND == SUB_NONNEG;
    /* order by exponent: A gets the larger one.  Lanes of a wave walk
       neighbouring pixels whose orbits mostly agree, so a step that no lane
       needs (here: any swap at all) is skipped for the whole wave. */
    const uint64_t swm = lane_mask(b_in.e > a_in.e);
    uint32_t A[N], B[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        A[j] = a_in.m[j];
        B[j] = b_in.m[j];
    }
    if (swm != 0) {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            A[j] = sel_mask(swm, a_in.m[j], b_in.m[j]);
            B[j] = sel_mask(swm, b_in.m[j], a_in.m[j]);
        }
    }
    /* signs of A and of the term B: for SUB_NONNEG, a is + and b enters as - */
    const uint32_t sbin = SUBNN ? 1u : (b_in.s ^ flipb);
    const uint32_t sa = SUBNN ? sel_mask(swm, 0u, 1u) : sel_mask(swm, a_in.s, sbin);
    const uint32_t sb = SUBNN ? (sa ^ 1u) : sel_mask(swm, sbin, a_in.s);
    int32_t e = (int32_t)sel_mask(swm, (uint32_t)a_in.e, (uint32_t)b_in.e);
    const int32_t eb = (int32_t)sel_mask(swm, (uint32_t)b_in.e, (uint32_t)a_in.e);

    /* align B right by d bits into N+1 words X (X[0] = guard word) + sticky */
    uint32_t d = (uint32_t)(e - eb);
    if (d > 32u * (N + 1)) d = 32u * (N + 1);
    /* |B| < 2^(e-d) is below a quarter of A's last place when d >= p + 2
       (p <= 32N): A - B lies within that quarter of A, nearer A than any
       other p-bit value (below a power of two the spacing halves, still more
       than the quarter), so RN_p(A - B) = A.  Near the real axis wim2 is
       that small against wre2 in every iteration, and the squares'
       difference skips its alignment, carry chain and rounding.  Only the
       squares' difference, only from 6 words (the vote in every addition of
       the 4-word loop cost config 2 4 %) and only in the word-aligned
       kernels (WORD_ALIGNED: p = 32N), whose register budget it leaves as
       it was; the ragged 8- and 16-word kernels spilled more with it. */
    if constexpr (SUBNN && WORD_ALIGNED && N >= 6) {
        if (wave_all(d >= 32u * N + 2u)) {
#pragma unroll
            for (int j = 0; j < N; ++j) r.m[j] = A[j];
            r.e = e;
            r.s = sa;
            return;
        }
    }
    const uint32_t ws = d >> 5, bs = d & 31u;
    uint32_t X[N + 1];
    X[0] = funnel_r(B[0], 0u, bs);
#pragma unroll
    for (int j = 1; j < N; ++j) X[j] = funnel_r(B[j], B[j - 1], bs);
    X[N] = B[N - 1] >> bs;
    uint32_t S = 0;
    if (wave_any(ws != 0)) {
#pragma unroll
        for (int st = 0; st < stages(N + 1); ++st) {
            const int t = 1 << st;
            const bool c = (ws & (uint32_t)t) != 0;
            uint32_t lost = 0;
#pragma unroll
            for (int j = 0; j < t && j <= N; ++j) lost |= X[j];
#pragma unroll
            for (int j = 0; j <= N; ++j) {
                const uint32_t nv = (j + t <= N) ? X[j + t] : 0u;
                X[j] = c ? nv : X[j];
            }
            S |= c ? lost : 0u;
        }
        S = (uint32_t)(S != 0);
    }

    /* T = A*2^32 +/- (X + S): one carry chain, X complemented for a
       subtraction with carry-in 1 - S (A - X - S = A + ~X + 1 - S); for
       SUB_NONNEG the borrow chain A*2^32 - X - S itself. */
    const uint32_t sub = sa ^ sb;
    uint32_t T[N + 1];
    uint32_t c;
    if (SUBNN) {
        uint32_t bw;
        T[0] = sub_bb(0u, X[0], S, bw);
#pragma unroll
        for (int j = 1; j <= N; ++j) T[j] = sub_bb(A[j - 1], X[j], bw, bw);
        c = bw ^ 1u;        /* the carry out of the complemented form */
    } else {
        const uint32_t msk = 0u - sub;
        c = sub & (S ^ 1u);
        T[0] = add_cc(X[0] ^ msk, 0u, c, c);
#pragma unroll
        for (int j = 1; j <= N; ++j) T[j] = add_cc(A[j - 1], X[j] ^ msk, c, c);
    }

    /* same signs with a carry out: shift right one bit */
    const bool rs = !SUBNN && (sub == 0) & (c != 0);
    if (!SUBNN && wave_any(rs)) {
        S |= rs ? (T[0] & 1u) : 0u;
#pragma unroll
        for (int j = 0; j < N; ++j) T[j] = rs ? funnel_r(T[j + 1], T[j], 1u) : T[j];
        T[N] = rs ? ((T[N] >> 1) | 0x80000000u) : T[N];
        e += rs ? 1 : 0;
    }

    /* a borrow (only possible when d == 0, so S == 0): negate, flip sign */
    uint32_t sres = sa;
    /* 0/1 from the data (c is the carry out, 0 or 1), used as a number */
    const uint32_t ngv = (SUBNN ? 1u : sub) & (c ^ 1u);
    if (wave_any(ngv != 0)) {
        uint32_t cc = ngv;
        const uint32_t nm = 0u - ngv;
#pragma unroll
        for (int j = 0; j <= N; ++j) T[j] = add_cc(T[j] ^ nm, 0u, cc, cc);
        sres ^= ngv;
    }

    /* cancellation of the whole top word (a subtraction with d <= 1, which
       is exact: S == 0): word shift */
    bool zero = false;
    if (__builtin_expect(wave_any(T[N] == 0), 0)) {
        uint32_t lzw = 0;
        bool found = false;
#pragma unroll
        for (int j = N; j >= 0; --j) {
            const bool here = !found && (T[j] != 0);
            lzw = here ? (uint32_t)(N - j) : lzw;
            found = found || (T[j] != 0);
        }
        zero = !found;
#pragma unroll
        for (int st = 0; st < stages(N); ++st) {
            const int t = 1 << st;
            const bool cnd = (lzw & (uint32_t)t) != 0;
#pragma unroll
            for (int j = N; j >= 0; --j) {
                const uint32_t nv = (j - t >= 0) ? T[j - t] : 0u;
                T[j] = cnd ? nv : T[j];
            }
        }
        e -= (int32_t)(32u * lzw);
    }

    /* bit normalisation: left shift by the leading zeros of the top word
       (0 or 1 unless the subtraction was exact) */
    const uint32_t lz = clz32(T[N]) & 31u;
    if (wave_any(lz != 0)) {
#pragma unroll
        for (int j = N; j >= 1; --j) T[j] = lz ? funnel_r(T[j], T[j - 1], 32u - lz) : T[j];
        T[0] = T[0] << lz;
        e -= (int32_t)lz;
    }

    /* an exact zero leaves T all zero words through the shifts and the
       rounding (no guard, no sticky), so only exponent and sign change */
    round_p<N>(T + 1, T[0], S, e, kclr);
#pragma unroll
    for (int j = 0; j < N; ++j) r.m[j] = T[j + 1];
    r.e = zero ? EZERO : e;
    r.s = zero ? 0u : sres;
}

/* r = 2a (exact) */
template <int N>
MP_HD void twice(mp<N>& a)
{
    a.e += is_zero(a) ? 0 : 1;
}

/* a > 4 (a normalised or zero) */
template <int N>
MP_HD bool gt4(const mp<N>& a)
{
    uint32_t low = 0;
#pragma unroll
    for (int j = 0; j < N - 1; ++j) low |= a.m[j];
    const bool above = (a.m[N - 1] != 0x80000000u) | (low != 0);
    return (a.s == 0) & !is_zero(a) & ((a.e > 3) | ((a.e == 3) & above));
}

/* value of a nonnegative a from its top word: a * (1 - 2^-31) < approx <= a */
template <int N>
MP_HD double approx_nonneg(const mp<N>& a)
{
    return ldexp((double)a.m[N - 1], a.e - 32);
}

/* RN_p(wim2 + wre2) > 4 for nonnegative wim2, wre2: decided from the top
   words unless the sum lies within 2^-26 of the threshold 4 + 2^(2-p), in
   which case the exact sum is rounded (src/frac_mandel.c:47-48).  Outside
   that window the top words decide: they fall short of the sum by less than
   2^-31 of it (under 2^-28 for sums below 8), and the double additions add
   less than 2^-50. */
template <int N>
MP_HD bool escaped(const mp<N>& wim2, const mp<N>& wre2, double thresh, int kclr)
{
    /* both squares below 2 (exponent <= 1) on every lane of the wave: sums
       below 4, no escape, and no estimate to form (most iterations of a bounded
       orbit take this exit) */
    if (wave_all((wim2.e <= 1) & (wre2.e <= 1)))
        return false;
    const double sa = approx_nonneg<N>(wim2) + approx_nonneg<N>(wre2);
    const double dlt = sa - thresh;
    bool esc = dlt > 0.0;
    const bool unsure = fabs(dlt) <= 0x1p-26;
    if (__builtin_expect(wave_any(unsure), 0)) {     /* out of line, as in mul_rn */
        mp<N> t;
        add_rn<N>(t, wim2, wre2, 0u, kclr);
        esc = unsure ? gt4<N>(t) : esc;
    }
    return esc;
}

} /* namespace mdzg */

#endif
