/*
 * mpf_gmp.h -- GMP's truncating mpf floats for the GMP path of the escape
 * loop (src/frac_mandel.c:55-80, src/frac_burning_ship.c:58-83,
 * src/frac_generalized_celtic.c:58-83, src/frac_variant.c:58-87), host and
 * device.
 *
 * MDZ's GMP path uses mpf_mul, mpf_mul_ui(.., 2), mpf_add, mpf_sub, mpf_abs
 * and mpf_cmp of GMP 6.2.1 (the library MDZ links; it is not part of the
 * reference tree).  These operations do not round: every operand and result
 * is a string of 64-bit limbs, at most PREC + 1 of them, where
 * PREC = (max(53, p) + 127) / 64 for mpf_init2(x, p), and each operation
 * keeps a limb-granular window of its operands and of its exact result.
 * This file restates those windows (from GMP's documented mpf semantics and
 * its mpf/add.c, sub.c, mul.c, mul_ui.c, abs.c and cmp.c algorithms); the
 * host tests compare every operation with libgmp itself, limb for limb.
 *
 * Representation (struct gf<L>, L >= PREC + 1 limbs of storage; the
 * operations take lr = PREC + 1 at run time): t[0] is the most significant
 * limb, t[i] for i < size are the limbs GMP holds (low zero limbs included:
 * GMP's size counts them), t[i] = 0 beyond; value = (-1)^neg *
 * sum t[i] * 2^(64 (exp - 1 - i)).  A nonzero value has t[0] != 0, as in GMP.
 */
#ifndef MDZG_MPF_GMP_H
#define MDZG_MPF_GMP_H

#include <stdint.h>

#include "mp_rn.h"   /* the 32 x 32 + 96-bit multiply-accumulates */

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#define GF_HD __host__ __device__ __forceinline__
#else
#define GF_HD static inline
#endif

namespace mdzg {
namespace gmpf {

typedef uint64_t limb;
static constexpr limb LMAX = ~(limb)0;

template <int L>
struct gf {
    limb t[L];
    int32_t size;
    int32_t exp;
    uint32_t neg;
};

/* a + b + cin (cin 0 or 1), carry to cout.  On the device a limb chain is
   one 32-bit v_add_co / v_addc_co chain (two instructions a limb) rather
   than 64-bit adds with compares for the carries. */
GF_HD limb addc(limb a, limb b, limb cin, limb& cout)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned int c0, c1;
    const uint32_t lo = __builtin_addc((uint32_t)a, (uint32_t)b, (uint32_t)cin, &c0);
    const uint32_t hi = __builtin_addc((uint32_t)(a >> 32), (uint32_t)(b >> 32), c0, &c1);
    cout = c1;
    return ((limb)hi << 32) | lo;
#else
    const limb s = a + b;
    const limb c1 = (limb)(s < a);
    const limb r = s + cin;
    cout = c1 | (limb)(r < s);
    return r;
#endif
}

/* a - b - bin (bin 0 or 1), borrow to bout (v_sub_co / v_subb_co) */
GF_HD limb subb(limb a, limb b, limb bin, limb& bout)
{
#if defined(__HIP_DEVICE_COMPILE__)
    unsigned int b0, b1;
    const uint32_t lo = __builtin_subc((uint32_t)a, (uint32_t)b, (uint32_t)bin, &b0);
    const uint32_t hi = __builtin_subc((uint32_t)(a >> 32), (uint32_t)(b >> 32), b0, &b1);
    bout = b1;
    return ((limb)hi << 32) | lo;
#else
    const limb d = a - b;
    const limb b1 = (limb)(a < b);
    const limb r = d - bin;
    bout = b1 | (limb)(d < bin);
    return r;
#endif
}

GF_HD void mul64(limb a, limb b, limb& hi, limb& lo)
{
#if defined(__HIP_DEVICE_COMPILE__)
    lo = a * b;
    hi = __umul64hi(a, b);
#else
    const unsigned __int128 p = (unsigned __int128)a * b;
    lo = (limb)p;
    hi = (limb)(p >> 64);
#endif
}

template <int L>
GF_HD void gf_zero(gf<L>& r)
{
#pragma unroll
    for (int i = 0; i < L; ++i) r.t[i] = 0;
    r.size = 0;
    r.exp = 0;
    r.neg = 0;
}

/* true when the predicate holds on some lane of the wave (device) / on
   this value (host): products that are zero on every lane are skipped */
GF_HD bool gf_any(bool c)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_ballot_w64(c) != 0;
#else
    return c;
#endif
}

/* ---- fixed-index array helpers (the device keeps the arrays in registers:
   every index below is a compile-time constant, run-time quantities only
   select values) --------------------------------------------------------- */

/* a[i] = 0 for i >= n */
template <int L>
GF_HD void keep_top(limb (&a)[L], int n)
{
#pragma unroll
    for (int i = 0; i < L; ++i) a[i] = (i < n) ? a[i] : (limb)0;
}

/* shift toward the less significant end by k limbs (k >= 0), zeros enter at
   the top and limbs pushed past position L-1 are lost.  Exponent gaps are
   mostly 0 or 1 limb, so a stage no lane of the wave needs is skipped. */
template <int L>
GF_HD void shift_down(limb (&a)[L], int k)
{
#pragma unroll
    for (int st = 1; st < 2 * L; st <<= 1) {
        const bool c = (k & st) != 0;
        if (gf_any(c)) {
#pragma unroll
            for (int i = L - 1; i >= 0; --i) a[i] = c ? ((i >= st) ? a[i - st] : (limb)0) : a[i];
        }
    }
    if (k >= 2 * L) {
#pragma unroll
        for (int i = 0; i < L; ++i) a[i] = 0;
    }
}

/* drop the top limb: a[i] = a[i+1] */
template <int L>
GF_HD void drop_top(limb (&a)[L])
{
#pragma unroll
    for (int i = 0; i < L - 1; ++i) a[i] = a[i + 1];
    a[L - 1] = 0;
}

/* a[i] for a run-time i in [0, L) (0 otherwise) */
template <int L>
GF_HD limb pick(const limb (&a)[L], int i)
{
    limb x = 0;
#pragma unroll
    for (int k = 0; k < L; ++k) x = (k == i) ? a[k] : x;
    return x;
}

/* columns Q .. 2W-2 of the product of the W-word numbers UW and VW (least
   significant word first) into ZW, each column's products added in
   multiply-accumulate groups (mp_rn.h mac_col) to the 96-bit accumulator
   that carries the column below */
template <int W, int Q>
GF_HD void word_cols(const uint32_t* UW, const uint32_t* VW, uint32_t* ZW, uint64_t& acc)
{
    if constexpr (Q < 2 * W - 1) {
        constexpr int i0 = Q < W ? 0 : Q - W + 1;
        constexpr int n = Q < W ? Q + 1 : 2 * W - 1 - Q;
        uint32_t ca[n], cb[n];
#pragma unroll
        for (int t = 0; t < n; ++t) {
            ca[t] = UW[i0 + t];
            cb[t] = VW[Q - i0 - t];
        }
        uint32_t c2 = 0;
        if constexpr (Q == 2 * W - 2)
            acc += (uint64_t)ca[0] * cb[0];   /* top column: the sum fits 64 bits */
        else
            mdzg::mac_col<n, true>(ca, cb, acc, c2);
        ZW[Q] = (uint32_t)acc;
        acc = (acc >> 32) | ((uint64_t)c2 << 32);
        word_cols<W, Q + 1>(UW, VW, ZW, acc);
    }
}

/* Z (2L limbs, most significant first) = U * V, where U and V (L limbs,
   most significant first) are zero below their top K limbs: the 2K-word
   numbers of those limbs (32-bit words, least significant first) multiplied
   column by column with a 96-bit accumulator, one v_mad_u64_u32 + v_addc per
   word product; the 2(L - K) low limbs of Z are zero. */
template <int L, int K>
GF_HD void word_product(const limb (&U)[L], const limb (&V)[L], limb (&Z)[2 * L])
{
    constexpr int W = 2 * K;
    uint32_t UW[W], VW[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int li = K - 1 - (k >> 1);
        UW[k] = (k & 1) ? (uint32_t)(U[li] >> 32) : (uint32_t)U[li];
        VW[k] = (k & 1) ? (uint32_t)(V[li] >> 32) : (uint32_t)V[li];
    }
    uint32_t ZW[2 * W];
    uint64_t acc = 0;
    word_cols<W, 0>(UW, VW, ZW, acc);
    ZW[2 * W - 1] = (uint32_t)acc;
#pragma unroll
    for (int m = 0; m < 2 * L; ++m) {
        const int k = 2 * (2 * K - 1 - m);
        Z[m] = (m < 2 * K) ? ((limb)ZW[k] | ((limb)ZW[k + 1] << 32)) : (limb)0;
    }
}

/* words Q .. 2W-1 of the square of the W-word number UW (least significant
   word first) into ZW: column Q of the off-diagonal sum O = sum over i < j of
   UW[i] UW[j] 2^(32(i+j)) is product-scanned in multiply-accumulate groups;
   once its word o is final, word Q of the square, 2 O + the diagonal
   squares, follows from o, the previous word of O, one word of a diagonal
   square and the carry c of that pass (mp_rn.h sqr_cols does the same for
   the top words of a rounded square) */
template <int W, int Q>
GF_HD void word_sqr_cols(const uint32_t* UW, uint32_t* ZW, uint64_t& acc, uint32_t& prev,
                         uint32_t& c, uint32_t& dhi)
{
    if constexpr (Q < 2 * W) {
        uint32_t o;
        if constexpr (Q <= 2 * W - 3) {
            constexpr int i0 = Q < W ? 0 : Q - W + 1;
            constexpr int i1 = (Q + 1) / 2 - 1;     /* largest i with i < Q - i */
            constexpr int n = i1 >= i0 ? i1 - i0 + 1 : 0;
            uint32_t c2 = 0;
            if constexpr (n > 0) {
                uint32_t ca[n], cb[n];
#pragma unroll
                for (int t = 0; t < n; ++t) {
                    ca[t] = UW[i0 + t];
                    cb[t] = UW[Q - i0 - t];
                }
                mdzg::mac_col<n, true>(ca, cb, acc, c2);
            }
            o = (uint32_t)acc;
            acc = (acc >> 32) | ((uint64_t)c2 << 32);
        } else {
            o = (Q == 2 * W - 2) ? (uint32_t)acc : (uint32_t)(acc >> 32);
        }
        uint32_t d;
        if constexpr ((Q & 1) == 0) {
            const uint64_t sq = (uint64_t)UW[Q >> 1] * UW[Q >> 1];
            d = (uint32_t)sq;
            dhi = (uint32_t)(sq >> 32);
        } else {
            d = dhi;
        }
        ZW[Q] = mdzg::add_cc(mdzg::funnel_r(o, prev, 31u), d, c, c);
        prev = o;
        word_sqr_cols<W, Q + 1>(UW, ZW, acc, prev, c, dhi);
    }
}

/* Z = U * U as word_product, with each off-diagonal word product formed
   once */
template <int L, int K>
GF_HD void word_square(const limb (&U)[L], limb (&Z)[2 * L])
{
    constexpr int W = 2 * K;
    uint32_t UW[W];
#pragma unroll
    for (int k = 0; k < W; ++k) {
        const int li = K - 1 - (k >> 1);
        UW[k] = (k & 1) ? (uint32_t)(U[li] >> 32) : (uint32_t)U[li];
    }
    uint32_t ZW[2 * W];
    uint64_t acc = 0;
    uint32_t prev = 0, c = 0, dhi = 0;
    word_sqr_cols<W, 0>(UW, ZW, acc, prev, c, dhi);
#pragma unroll
    for (int m = 0; m < 2 * L; ++m) {
        const int k = 2 * (2 * K - 1 - m);
        Z[m] = (m < 2 * K) ? ((limb)ZW[k] | ((limb)ZW[k + 1] << 32)) : (limb)0;
    }
}

/* ------------------------------------------------------------------ */
/* mpf_mul (mpf/mul.c): each operand cut to its top PREC limbs, the exact
   product of those, a zero top limb dropped, the top PREC + 1 limbs kept.   */
/* ------------------------------------------------------------------ */
template <int L, bool SQR = false>
GF_HD void gf_mul(gf<L>& r, const gf<L>& u, const gf<L>& v, int lr)
{
    const int P = lr - 1;
    const int us = u.size < P ? u.size : P;
    const int vs = v.size < P ? v.size : P;
    /* Limbs beyond a value's size are zero.  So when every lane of the wave
       runs at the bucket's own PREC + 1 = L limbs, the cut to PREC limbs
       only clears the last limb, and the product's limbs from us + vs on are
       zero, which leaves the result window without selects. */
    const bool full = !gf_any(lr != L);
    limb U[L], V[L];
    if (full) {
#pragma unroll
        for (int i = 0; i < L - 1; ++i) {
            U[i] = u.t[i];
            V[i] = v.t[i];
        }
        U[L - 1] = 0;
        V[L - 1] = 0;
    } else {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            U[i] = (i < us) ? u.t[i] : (limb)0;
            V[i] = (i < vs) ? v.t[i] : (limb)0;
        }
    }
    /* U and V hold at most PREC <= L - 1 limbs, so their last limb is zero;
       the product of their top K limbs is formed by word_product.  K = L - 1
       always suffices; the small buckets, which often hold fewer limbs than
       they store, take K = L - 2 when no lane of the wave needs more. */
    limb Z[2 * L];
    /* squares by word_square from 6 words (4-limb bucket: +4 %, 6-limb
       bucket: +7 % on the config 2 view against the plain product) */
    const bool small = L <= 6 && L >= 3 && !gf_any((us > L - 2) | (vs > L - 2));
    if (SQR && 2 * (L - 1) >= 6 && small)
        word_square<L, (L >= 3 ? L - 2 : 1)>(U, Z);
    else if (SQR && 2 * (L - 1) >= 6)
        word_square<L, L - 1>(U, Z);
    else if (small)
        word_product<L, (L >= 3 ? L - 2 : 1)>(U, V, Z);
    else
        word_product<L, L - 1>(U, V, Z);
    const bool zero = (us == 0) | (vs == 0);
    const int adj = (Z[0] == 0) ? 1 : 0;
    const int rsize = us + vs - adj;
    const int keep = zero ? 0 : (rsize < lr ? rsize : lr);
    if (full) {     /* a zero operand leaves Z zero: no case for it either */
#pragma unroll
        for (int k = 0; k < L; ++k) r.t[k] = adj ? Z[k + 1] : Z[k];
    } else {
#pragma unroll
        for (int k = 0; k < L; ++k) r.t[k] = (k < keep) ? (adj ? Z[k + 1] : Z[k]) : (limb)0;
    }
    r.size = keep;
    r.exp = zero ? 0 : u.exp + v.exp - adj;
    r.neg = zero ? 0 : (u.neg ^ v.neg);
}

/* mpf_mul (x, u, u) */
template <int L>
GF_HD void gf_sqr(gf<L>& r, const gf<L>& u, int lr)
{
    gf_mul<L, true>(r, u, u, lr);
}

/* ------------------------------------------------------------------ */
/* mpf_mul_ui (mpf/mul_ui.c) by 2: the operand cut to its top PREC limbs,
   with the carry the discarded limbs would have produced; a carry out of
   the top adds a limb.                                                      */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_mul2(gf<L>& r, const gf<L>& u, int lr)
{
    const int P = lr - 1;
    int size = u.size;
    limb tmp[L];
    limb c;
    if (!gf_any(lr != L)) {
        /* PREC = L - 1 on every lane: the cut drops t[L-1] (zero unless
           size = L), and zero limbs shift in as zeros with no carry */
        size = size < P ? size : P;
        c = u.t[L - 1] >> 63;
        tmp[L - 1] = 0;
#pragma unroll
        for (int i = L - 2; i >= 0; --i) {
            tmp[i] = (u.t[i] << 1) | c;
            c = u.t[i] >> 63;
        }
    } else {
        limb cin = 0;
        if (size > P) {
            cin = pick<L>(u.t, P) >> 63;
            size = P;
        }
        c = cin;
#pragma unroll
        for (int i = L - 1; i >= 0; --i) {
            const bool in = i < size;
            tmp[i] = in ? ((u.t[i] << 1) | c) : (limb)0;
            c = in ? (u.t[i] >> 63) : c;
        }
    }
    const bool grow = c != 0;
#pragma unroll
    for (int i = L - 1; i >= 1; --i) r.t[i] = grow ? tmp[i - 1] : tmp[i];
    r.t[0] = grow ? (limb)1 : tmp[0];
    r.size = u.size ? size + (grow ? 1 : 0) : 0;
    r.exp = u.size ? u.exp + (grow ? 1 : 0) : 0;
    r.neg = u.size ? u.neg : 0;
}

/* mpf_abs in place */
template <int L>
GF_HD void gf_abs(gf<L>& r)
{
    r.neg = 0;
}

/* mpf_cmp (x, 4) > 0 for the bailout value set by mpf_set_si (bail, 4) */
template <int L>
GF_HD bool gf_gt4(const gf<L>& x)
{
    limb rest = 0;
#pragma unroll
    for (int i = 1; i < L; ++i) rest |= x.t[i];           /* zero beyond size */
    const bool above = (x.exp > 1) | ((x.exp == 1) & ((x.t[0] > 4) | ((x.t[0] == 4) & (rest != 0))));
    return (x.size != 0) & (x.neg == 0) & above;
}

template <int L>
GF_HD void gf_sub_mag(gf<L>& r, const gf<L>& u0, const gf<L>& v0, uint32_t neg0, int lr);

/* ------------------------------------------------------------------ */
/* mpf_add (mpf/add.c)                                                   */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_add(gf<L>& r, const gf<L>& a, const gf<L>& b, uint32_t flipb, int lr)
{
    const int P = lr - 1;
    const uint32_t bneg = b.neg ^ flipb;
    if (a.size == 0) {
        r = b;
        r.neg = b.size ? bneg : 0;
        return;
    }
    if (b.size == 0) {
        r = a;
        return;
    }
    if (a.neg != bneg) {
        /* a + b with opposite signs: |a| - |b| with a's sign */
        gf_sub_mag<L>(r, a, b, a.neg, lr);
        return;
    }
    const bool sw = a.exp < b.exp;
    limb U[L], V[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        U[i] = sw ? b.t[i] : a.t[i];
        V[i] = sw ? a.t[i] : b.t[i];
    }
    const int uexp = sw ? b.exp : a.exp;
    const int ediff = sw ? b.exp - a.exp : a.exp - b.exp;
    int us = sw ? b.size : a.size, vs = sw ? a.size : b.size;
    if (us > P) us = P;
    if (vs + ediff > P) vs = P - ediff;
    keep_top<L>(U, us);
    int rsize;
    if (ediff >= P) {
        rsize = us;
#pragma unroll
        for (int i = 0; i < L; ++i) V[i] = 0;
    } else {
        keep_top<L>(V, vs);
        shift_down<L>(V, ediff);
        rsize = (us > ediff) ? ((vs + ediff <= us) ? us : vs + ediff) : vs + ediff;
    }
    limb R[L];
    limb c = 0;
#pragma unroll
    for (int k = L - 1; k >= 0; --k) R[k] = addc(U[k], V[k], c, c);
    const bool grow = c != 0;
#pragma unroll
    for (int k = L - 1; k >= 1; --k) r.t[k] = grow ? R[k - 1] : R[k];
    r.t[0] = grow ? (limb)1 : R[0];
    rsize += grow ? 1 : 0;
    keep_top<L>(r.t, rsize);
    r.size = rsize;
    r.exp = rsize ? uexp + (grow ? 1 : 0) : 0;
    r.neg = rsize ? a.neg : 0;
}

/* ------------------------------------------------------------------ */
/* mpf_sub (mpf/sub.c) on magnitudes: r = sign(neg0) * (|u0| - |v0|).      */
/* The windows of u and v are moved to the top of U and V as their leading */
/* limbs are consumed.                                                      */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_sub_mag(gf<L>& r, const gf<L>& u0, const gf<L>& v0, uint32_t neg0, int lr)
{
    const int prec = lr;         /* PREC(r) + 1 */
    const bool sw = u0.exp < v0.exp;
    uint32_t negate = neg0 ^ (sw ? 1u : 0u);
    limb U[L], V[L], T[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        U[i] = sw ? v0.t[i] : u0.t[i];
        V[i] = sw ? u0.t[i] : v0.t[i];
        T[i] = 0;
    }
    int usize = sw ? v0.size : u0.size, vsize = sw ? u0.size : v0.size;
    int exp = sw ? v0.exp : u0.exp;
    const int ediff = sw ? v0.exp - u0.exp : u0.exp - v0.exp;
    int rsize = 0;
    int mode = 0;      /* 0 general, 1 near-cancellation result, 2 copy of V, 3 copy of U */
    bool norm = true;
    bool one = false;  /* a leading limb 1 above T (near-cancellation, no borrow) */

    if (ediff == 0) {
        bool done = false;
        while (!done && U[0] == V[0]) {
            drop_top<L>(U);
            drop_top<L>(V);
            --usize;
            --vsize;
            --exp;
            if (usize == 0 || vsize == 0) {
                /* one operand cancels the other's top limbs: the rest of the
                   longer one, high zero limbs stripped, cut to prec */
                if (usize == 0) {
                    negate ^= 1;
                } else {
#pragma unroll
                    for (int i = 0; i < L; ++i) V[i] = U[i];
                    vsize = usize;
                }
                while (vsize != 0 && V[0] == 0) {
                    drop_top<L>(V);
                    --vsize;
                    --exp;
                }
                if (vsize > prec) vsize = prec;
                mode = 2;
                done = true;
            }
        }
        if (mode == 0) {
            if (U[0] < V[0]) {
#pragma unroll
                for (int i = 0; i < L; ++i) {
                    const limb t = U[i];
                    U[i] = V[i];
                    V[i] = t;
                }
                const int t = usize;
                usize = vsize;
                vsize = t;
                negate ^= 1;
            }
            if (U[0] == V[0] + 1) {
                drop_top<L>(U);
                drop_top<L>(V);
                --usize;
                --vsize;
                --exp;
                mode = 1;
            }
        }
    } else if (ediff == 1) {
        if (U[0] == 1 && V[0] == LMAX && !(usize >= 2 && U[1] != 0)) {
            drop_top<L>(U);
            --usize;
            --exp;
            mode = 1;
        }
    }

    if (mode == 1) {
        while (vsize != 0 && usize != 0 && U[0] == 0 && V[0] == LMAX) {
            drop_top<L>(U);
            drop_top<L>(V);
            --usize;
            --vsize;
            --exp;
        }
        if (usize == 0) {
            while (vsize != 0 && V[0] == LMAX) {
                drop_top<L>(V);
                --vsize;
                --exp;
            }
        } else if (usize > prec - 1) {
            usize = prec - 1;
        }
        if (vsize > prec - 1) vsize = prec - 1;
        keep_top<L>(U, usize);
        keep_top<L>(V, vsize);
        if (vsize == 0) {
            /* 1 followed by u's limbs */
#pragma unroll
            for (int k = 0; k < L; ++k) T[k] = U[k];
            rsize = usize;
            one = true;
            norm = false;
        } else {
            /* (1 . u) - (0 . v) on the aligned windows, both cut to prec - 1
               limbs: their difference, and no borrow keeps the leading 1 */
            rsize = usize > vsize ? usize : vsize;
            limb bw = 0;
#pragma unroll
            for (int k = L - 1; k >= 0; --k) T[k] = subb(U[k], V[k], bw, bw);
            one = (bw == 0);
        }
    } else if (mode == 0) {
        /* general case */
        if (usize > prec) usize = prec;
        if (vsize + ediff > prec) vsize = prec - ediff;
        if (ediff >= prec) {
            mode = 3;
            norm = false;
        } else {
            keep_top<L>(U, usize);
            keep_top<L>(V, vsize);
            while (vsize > 0 && pick<L>(V, vsize - 1) == 0) --vsize;
            if (vsize <= 0) {
                mode = 3;
                norm = false;
            } else {
                while (usize > 0 && pick<L>(U, usize - 1) == 0) --usize;
                /* usize == 0 cannot happen: U's top limb is nonzero */
                shift_down<L>(V, ediff);
                rsize = (usize > ediff) ? ((vsize + ediff <= usize) ? usize : vsize + ediff)
                                        : vsize + ediff;
                limb bw = 0;
#pragma unroll
                for (int k = L - 1; k >= 0; --k) T[k] = subb(U[k], V[k], bw, bw);
            }
        }
    }

    if (mode == 2) {
#pragma unroll
        for (int k = 0; k < L; ++k) T[k] = V[k];
        rsize = vsize;
        norm = false;
    } else if (mode == 3) {
#pragma unroll
        for (int k = 0; k < L; ++k) T[k] = U[k];
        rsize = usize;
    }
    if (one) {
        /* the leading 1 goes on top */
#pragma unroll
        for (int k = L - 1; k >= 1; --k) T[k] = T[k - 1];
        T[0] = 1;
        ++rsize;
        ++exp;
    }
    keep_top<L>(T, rsize);
    if (norm) {
        while (rsize > 0 && T[0] == 0) {
            drop_top<L>(T);
            --rsize;
            --exp;
        }
    }
#pragma unroll
    for (int k = 0; k < L; ++k) r.t[k] = T[k];
    r.size = rsize;
    r.exp = rsize ? exp : 0;
    r.neg = rsize ? negate : 0;
}

/* r = a - b: signs as mpf_sub, which adds when they differ */
template <int L>
GF_HD void gf_sub(gf<L>& r, const gf<L>& a, const gf<L>& b, int lr)
{
    if (a.size == 0) {
        r = b;
        r.neg = b.size ? (b.neg ^ 1u) : 0;
        return;
    }
    if (b.size == 0) {
        r = a;
        return;
    }
    if (a.neg != b.neg) {
        gf_add<L>(r, a, b, 1u, lr);
        return;
    }
    gf_sub_mag<L>(r, a, b, a.neg, lr);
}

/* ------------------------------------------------------------------ */
/* r = a + (-1)^flipb b: mpf_add and mpf_sub in one carry chain.         */
/*                                                                        */
/* The common case of both -- nonzero operands and, for a subtraction of  */
/* magnitudes, neither equal leading limbs at equal exponents nor one of  */
/* mpf_sub's two near-cancellation forms -- has one shape: U (the operand */
/* of larger exponent; of larger leading limb at equal exponents when     */
/* subtracting) cut to `lim` limbs (PREC for an addition, PREC + 1 for a  */
/* subtraction), V cut to the same window, for a subtraction both         */
/* stripped of low zero limbs, V shifted down by the exponent gap, then   */
/* U + V, or U + ~V + 1 = U - V.  An addition's carry out adds a leading  */
/* 1 limb; a subtraction loses at most one leading limb (U's top limb is  */
/* at least one unit above V's contribution outside the excluded forms).  */
/* Lanes outside the common case take gf_add's general code, which a     */
/* wave runs only when one of its lanes needs it.                         */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_addsub(gf<L>& r, const gf<L>& a, const gf<L>& b, uint32_t flipb, int lr)
{
    const uint32_t bneg = b.neg ^ flipb;
    const bool sub = a.neg != bneg;
    const bool sw = (a.exp < b.exp) | (sub & (a.exp == b.exp) & (a.t[0] < b.t[0]));
    /* values first: selects between the operands' fields through their
       addresses would keep the operands out of registers */
    limb A[L], B[L], U[L], V[L];
#pragma unroll
    for (int i = 0; i < L; ++i) {
        A[i] = a.t[i];
        B[i] = b.t[i];
        U[i] = A[i];
        V[i] = B[i];
    }
    if (gf_any(sw)) {
#pragma unroll
        for (int i = 0; i < L; ++i) {
            U[i] = sw ? B[i] : A[i];
            V[i] = sw ? A[i] : B[i];
        }
    }
    const int uexp = sw ? b.exp : a.exp;
    const int ediff = sw ? b.exp - a.exp : a.exp - b.exp;
    int us = sw ? b.size : a.size, vs = sw ? a.size : b.size;
    const bool near0 = (ediff == 0) & ((U[0] == V[0]) | (U[0] == V[0] + 1));
    const bool near1 = (ediff == 1) & (U[0] == 1) & (V[0] == LMAX) & !((us >= 2) & (U[1] != 0));
    const bool fast = (a.size != 0) & (b.size != 0) & !(sub & (near0 | near1));

    const int lim = sub ? lr : lr - 1;
    if (us > lim) us = lim;
    if (vs + ediff > lim) vs = lim - ediff;      /* <= 0: V drops out */
    if (!gf_any(lr != L))
        U[L - 1] = (us < L) ? (limb)0 : U[L - 1];  /* zero beyond the size: only
                                                     a cut to L - 1 clears */
    else
        keep_top<L>(U, us);
    keep_top<L>(V, vs);
    int ulast = 0, vlast = 0;                    /* the windows without low zero limbs */
#pragma unroll
    for (int i = 0; i < L; ++i) {
        ulast = (U[i] != 0) ? i + 1 : ulast;
        vlast = (V[i] != 0) ? i + 1 : vlast;
    }
    const int ue = sub ? ulast : us;
    const int ve = sub ? vlast : vs;
    int rsize = (ve <= 0) ? us : ((ue > ediff) ? ((ve + ediff <= ue) ? ue : ve + ediff) : ve + ediff);
    shift_down<L>(V, ediff);
    limb T[L];
    const limb m = sub ? LMAX : (limb)0;
    limb c = sub ? 1 : 0;
#pragma unroll
    for (int k = L - 1; k >= 0; --k) T[k] = addc(U[k], V[k] ^ m, c, c);
    int exp = uexp;
    const bool grow = !sub & (c != 0);
    if (gf_any(grow)) {
#pragma unroll
        for (int k = L - 1; k >= 1; --k) T[k] = grow ? T[k - 1] : T[k];
        T[0] = grow ? (limb)1 : T[0];
        rsize += grow ? 1 : 0;
        exp += grow ? 1 : 0;
    }
    const bool nrm = sub & (T[0] == 0);
    if (gf_any(nrm)) {
#pragma unroll
        for (int k = 0; k < L - 1; ++k) T[k] = nrm ? T[k + 1] : T[k];
        T[L - 1] = nrm ? (limb)0 : T[L - 1];
        rsize -= nrm ? 1 : 0;
        exp -= nrm ? 1 : 0;
    }
    uint32_t neg = a.neg ^ ((sub & sw) ? 1u : 0u);
    if (__builtin_expect(gf_any(!fast), 0)) {    /* rare: laid out past the loop */
        gf<L> s;
        gf_add<L>(s, a, b, flipb, lr);
#pragma unroll
        for (int k = 0; k < L; ++k) T[k] = fast ? T[k] : s.t[k];
        rsize = fast ? rsize : s.size;
        exp = fast ? exp : s.exp;
        neg = fast ? neg : s.neg;
    }
    /* r may be a or b: written last */
#pragma unroll
    for (int k = 0; k < L; ++k) r.t[k] = T[k];
    r.size = rsize;
    r.exp = exp;
    r.neg = neg;
}

/* mpf_add (t, a, b); mpf_cmp (t, 4) > 0 for nonnegative a and b (the
   squares): 1 or 0 when the integer limbs decide it, -1 otherwise.  The sum
   of the operands' windows lies in [ia + ib, a + b] with ia, ib the integer
   parts (the top limbs at exponent 1, which every window keeps), and
   a + b < ia + ib + 2, so ia + ib <= 2 means no escape and ia + ib >= 5
   (or an operand of 2^64 or more) an escape. */
template <int L>
GF_HD int gf_sum_gt4_quick(const gf<L>& a, const gf<L>& b)
{
    const bool big = ((a.size != 0) & (a.exp >= 2)) | ((b.size != 0) & (b.exp >= 2));
    const limb ia = ((a.size != 0) & (a.exp == 1)) ? a.t[0] : (limb)0;
    const limb ib = ((b.size != 0) & (b.exp == 1)) ? b.t[0] : (limb)0;
    const bool high = big | (ia >= 5) | (ib >= 5) | (ia + ib >= 5);
    return high ? 1 : ((ia + ib <= 2) ? 0 : -1);
}

}  // namespace gmpf
}  // namespace mdzg

#endif
