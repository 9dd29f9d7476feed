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

GF_HD limb addc(limb a, limb b, limb cin, limb& cout)
{
    const limb s = a + b;
    const limb c1 = (limb)(s < a);
    const limb r = s + cin;
    cout = c1 | (limb)(r < s);
    return r;
}

GF_HD limb subb(limb a, limb b, limb bin, limb& bout)
{
    const limb d = a - b;
    const limb b1 = (limb)(a < b);
    const limb r = d - bin;
    bout = b1 | (limb)(d < bin);
    return r;
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
    for (int i = 0; i < L; ++i) r.t[i] = 0;
    r.size = 0;
    r.exp = 0;
    r.neg = 0;
}

/* limb i (from the top) of a window starting `off` limbs below the top of t,
   holding n limbs; 0 outside */
template <int L>
GF_HD limb wl(const limb* t, int off, int n, int i)
{
    return (i >= 0 && i < n && off + i < L) ? t[off + i] : (limb)0;
}

/* ------------------------------------------------------------------ */
/* mpf_mul (mpf/mul.c): each operand cut to its top PREC limbs, the exact
   product of those, a zero top limb dropped, the top PREC + 1 limbs kept.   */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_mul(gf<L>& r, const gf<L>& u, const gf<L>& v, int lr)
{
    const int P = lr - 1;
    const int us = u.size < P ? u.size : P;
    const int vs = v.size < P ? v.size : P;
    if (us == 0 || vs == 0) {
        gf_zero(r);
        return;
    }
    /* least significant first: U[i] = u.t[us-1-i] */
    limb W[2 * L];
    for (int k = 0; k < 2 * L; ++k) W[k] = 0;
    for (int i = 0; i < us; ++i) {
        const limb ui = u.t[us - 1 - i];
        limb carry = 0;
        for (int j = 0; j < vs; ++j) {
            limb hi, lo, c1, c2;
            mul64(ui, v.t[vs - 1 - j], hi, lo);
            lo = addc(lo, W[i + j], 0, c1);
            lo = addc(lo, carry, 0, c2);
            W[i + j] = lo;
            carry = hi + c1 + c2;
        }
        W[i + vs] = carry;
    }
    const int n = us + vs;
    const int adj = (W[n - 1] == 0) ? 1 : 0;
    const int rsize = n - adj;
    const int keep = rsize < lr ? rsize : lr;
    for (int k = 0; k < L; ++k) r.t[k] = (k < keep) ? W[n - 1 - adj - k] : 0;
    r.size = keep;
    r.exp = u.exp + v.exp - adj;
    r.neg = u.neg ^ v.neg;
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
    if (size == 0) {
        gf_zero(r);
        return;
    }
    limb cin = 0;
    if (size > P) {
        cin = u.t[P] >> 63;
        size = P;
    }
    limb tmp[L];
    limb c = cin;
    for (int i = L - 1; i >= 0; --i) {
        if (i < size) {
            tmp[i] = (u.t[i] << 1) | c;
            c = u.t[i] >> 63;
        } else {
            tmp[i] = 0;
        }
    }
    r.neg = u.neg;
    if (c) {
        r.t[0] = 1;
        for (int i = 1; i < L; ++i) r.t[i] = (i - 1 < size) ? tmp[i - 1] : 0;
        r.size = size + 1;
        r.exp = u.exp + 1;
    } else {
        for (int i = 0; i < L; ++i) r.t[i] = tmp[i];
        r.size = size;
        r.exp = u.exp;
    }
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
    if (x.size == 0 || x.neg)
        return false;
    if (x.exp != 1)
        return x.exp > 1;
    if (x.t[0] != 4)
        return x.t[0] > 4;
    limb rest = 0;
    for (int i = 1; i < L; ++i) rest |= (i < x.size) ? x.t[i] : 0;
    return rest != 0;
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
    const gf<L>& u = sw ? b : a;
    const gf<L>& v = sw ? a : b;
    const uint32_t negate = a.neg;
    const int ediff = u.exp - v.exp;
    int us = u.size, vs = v.size;
    if (us > P) us = P;
    if (vs + ediff > P) vs = P - ediff;
    limb R[L + 1];
    int rsize, exp = u.exp;
    if (ediff >= P) {
        for (int k = 0; k < L; ++k) R[k] = (k < us) ? u.t[k] : 0;
        rsize = us;
    } else {
        if (us > ediff)
            rsize = (vs + ediff <= us) ? us : vs + ediff;
        else
            rsize = vs + ediff;
        limb c = 0;
        for (int k = L - 1; k >= 0; --k) {
            const limb x = (k < us) ? u.t[k] : 0;
            const int j = k - ediff;
            const limb y = (j >= 0 && j < vs) ? v.t[j] : 0;
            R[k] = (k < rsize) ? addc(x, y, c, c) : 0;
            if (k >= rsize) c = 0;
        }
        if (c) {
            for (int k = L - 1; k >= 1; --k) R[k] = R[k - 1];
            R[0] = 1;
            ++rsize;
            ++exp;
        }
    }
    for (int k = 0; k < L; ++k) r.t[k] = (k < rsize) ? R[k] : 0;
    r.size = rsize;
    r.exp = rsize ? exp : 0;
    r.neg = rsize ? negate : 0;
}

/* ------------------------------------------------------------------ */
/* mpf_sub (mpf/sub.c) on magnitudes: r = sign(neg0) * (|u0| - |v0|).
   Windows of u and v are (offset from the top of t, limb count).          */
/* ------------------------------------------------------------------ */
template <int L>
GF_HD void gf_sub_mag(gf<L>& r, const gf<L>& u0, const gf<L>& v0, uint32_t neg0, int lr)
{
    const int prec = lr;         /* PREC(r) + 1 */
    uint32_t negate = neg0;
    const limb* up = u0.t;
    const limb* vp = v0.t;
    int uoff = 0, voff = 0;
    int usize = u0.size, vsize = v0.size;
    int exp = u0.exp;
    int ediff = u0.exp - v0.exp;
    if (u0.exp < v0.exp) {
        up = v0.t;
        vp = u0.t;
        usize = v0.size;
        vsize = u0.size;
        exp = v0.exp;
        ediff = -ediff;
        negate ^= 1;
    }
    limb T[L + 2];
    for (int k = 0; k < L + 2; ++k) T[k] = 0;
    int rsize = 0;
    bool norm = true;            /* strip leading zero limbs at the end */

#define U(i) wl<L>(up, uoff, usize, (i))
#define V(i) wl<L>(vp, voff, vsize, (i))
    if (ediff <= 1) {
        bool general = false;
        if (ediff == 0) {
            while (U(0) == V(0)) {
                ++uoff; --usize; ++voff; --vsize; --exp;
                if (usize == 0 || vsize == 0) {
                    /* one operand cancels the other's top limbs: the rest of
                       the longer one (high zero limbs stripped, cut to prec) */
                    if (usize == 0) {
                        negate ^= 1;
                    } else {
                        vp = up; voff = uoff; vsize = usize;
                    }
                    while (vsize != 0 && V(0) == 0) { ++voff; --vsize; --exp; }
                    if (vsize > prec) vsize = prec;
                    for (int k = 0; k < vsize; ++k) T[k] = V(k);
                    rsize = vsize;
                    goto done;
                }
            }
            if (U(0) < V(0)) {
                const limb* tp = up; up = vp; vp = tp;
                int t = uoff; uoff = voff; voff = t;
                t = usize; usize = vsize; vsize = t;
                negate ^= 1;
            }
            if (U(0) != V(0) + 1) {
                general = true;
            } else {
                ++uoff; --usize; ++voff; --vsize; --exp;
            }
        } else {
            if (U(0) != 1 || V(0) != LMAX || (usize >= 2 && U(1) != 0)) {
                general = true;
            } else {
                ++uoff; --usize; --exp;
            }
        }
        if (!general) {
            while (vsize != 0 && usize != 0 && U(0) == 0 && V(0) == LMAX) {
                ++uoff; --usize; ++voff; --vsize; --exp;
            }
            if (usize == 0) {
                while (vsize != 0 && V(0) == LMAX) { ++voff; --vsize; --exp; }
            } else if (usize > prec - 1) {
                usize = prec - 1;
            }
            if (vsize > prec - 1) vsize = prec - 1;
            if (vsize == 0) {
                /* 1 followed by u's limbs */
                T[0] = 1;
                for (int k = 0; k < usize; ++k) T[k + 1] = U(k);
                rsize = usize + 1;
                ++exp;
                norm = false;
                goto store;
            }
            /* exact (1 . u) - (0 . v) on the aligned windows, both cut to
               prec - 1 limbs: the difference of the windows with its borrow;
               no borrow keeps the leading 1 */
            const int n = usize > vsize ? usize : vsize;
            limb bw = 0;
            for (int k = n - 1; k >= 0; --k) T[k + 1] = subb(U(k), V(k), bw, bw);
            rsize = n;
            if (bw == 0) {
                T[0] = 1;
                ++rsize;
                ++exp;
            } else {
                for (int k = 0; k < n; ++k) T[k] = T[k + 1];
                T[n] = 0;
            }
            goto store;
        }
    }

    /* general case */
    if (usize > prec) usize = prec;
    if (vsize + ediff > prec) vsize = prec - ediff;
    if (ediff >= prec) {
        for (int k = 0; k < usize; ++k) T[k] = U(k);
        rsize = usize;
        norm = false;
        goto store;
    }
    while (true) {
        if (vsize <= 0) {
            for (int k = 0; k < usize; ++k) T[k] = U(k);
            rsize = usize;
            goto done;
        }
        if (V(vsize - 1) != 0) break;
        --vsize;
    }
    while (true) {
        if (usize == 0) {
            for (int k = 0; k < vsize; ++k) T[k] = V(k);
            rsize = vsize;
            negate ^= 1;
            goto done;
        }
        if (U(usize - 1) != 0) break;
        --usize;
    }
    {
        /* |u| - |v| with v ediff limbs lower, exact over the union span */
        rsize = (usize > ediff) ? ((vsize + ediff <= usize) ? usize : vsize + ediff)
                                : vsize + ediff;
        limb bw = 0;
        for (int k = rsize - 1; k >= 0; --k) T[k] = subb(U(k), V(k - ediff), bw, bw);
    }

store:
    if (norm) {
        int lead = 0;
        while (lead < rsize && T[lead] == 0) ++lead;
        for (int k = 0; k + lead < L + 2; ++k) T[k] = T[k + lead];
        rsize -= lead;
        exp -= lead;
    }
done:
#undef U
#undef V
    for (int k = 0; k < L; ++k) r.t[k] = (k < rsize) ? T[k] : 0;
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

}  // namespace gmpf
}  // namespace mdzg

#endif
