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
 * The header is plain C++ apart from the wave vote used to skip stages of the
 * variable shifters that no lane needs; built for the host (tests only) the
 * vote is the lane's own predicate, which skips exactly the same no-op work.
 */
#ifndef MDZG_MP_RN_H
#define MDZG_MP_RN_H

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
    return __any(c) != 0;
#else
    return c;
#endif
}

/* (hi:lo) >> s, low 32 bits, for 0 <= s < 32 (v_alignbit_b32) */
MP_HD uint32_t funnel_r(uint32_t hi, uint32_t lo, uint32_t s)
{
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> (s & 31u));
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
template <int N>
MP_HD void round_full(uint32_t* R, uint32_t G, uint32_t S, int32_t& e)
{
    uint32_t below = (G << 1) | S;
    uint32_t inc = (G >> 31) & (uint32_t)((below != 0) | (R[0] & 1u));
    uint32_t c = inc;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint32_t t = R[j] + c;
        c = (uint32_t)(t < c);
        R[j] = t;
    }
    if (c) {            /* R was all ones: now 2^(32N) -> 2^(32N-1), e+1 */
        R[N - 1] = 0x80000000u;
        e += 1;
    }
}

/* kclr = 32N - p in [1, 32N-1]: clear the low kclr bits of R, round at bit kclr */
template <int N>
MP_HD void round_masked(uint32_t* R, uint32_t G, uint32_t S, int32_t& e, int kclr)
{
    const int wi = kclr >> 5, bi = kclr & 31;          /* lsb of the kept part */
    const int rr = kclr - 1, rw = rr >> 5, rb = rr & 31; /* round bit          */
    uint32_t rbit = 0, st = G | S, lsb = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint32_t keep = (j > wi) ? 0xffffffffu : (j == wi ? (0xffffffffu << bi) : 0u);
        uint32_t rm = (j == rw) ? (1u << rb) : 0u;
        uint32_t sm = (j < rw) ? 0xffffffffu : (j == rw ? ((1u << rb) - 1u) : 0u);
        uint32_t lm = (j == wi) ? (1u << bi) : 0u;
        rbit |= R[j] & rm;
        st |= R[j] & sm;
        lsb |= R[j] & lm;
        R[j] &= keep;
    }
    uint32_t inc = (uint32_t)(rbit != 0) & (uint32_t)((st != 0) | (lsb != 0));
    uint32_t c = 0;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        uint32_t add = (j == wi) ? (inc << bi) : 0u;
        uint64_t t = (uint64_t)R[j] + add + c;
        R[j] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
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
/* Each column sum is kept in a 96-bit accumulator (c2:c1:c0).          */
/* ------------------------------------------------------------------ */
MP_HD void mac96(uint32_t a, uint32_t b, uint32_t& c0, uint32_t& c1, uint32_t& c2)
{
    uint64_t acc = ((uint64_t)c1 << 32) | c0;
    uint64_t p = (uint64_t)a * b;
    uint64_t t = acc + p;
    c2 += (uint32_t)(t < p);
    c0 = (uint32_t)t;
    c1 = (uint32_t)(t >> 32);
}

template <int N>
MP_HD void mant_mul(const uint32_t* a, const uint32_t* b, uint32_t* H, uint32_t& S)
{
    uint32_t c0 = 0, c1 = 0, c2 = 0, st = 0;
#pragma unroll
    for (int k = 0; k < 2 * N - 1; ++k) {
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int j = k - i;
            if (j >= 0 && j < N)
                mac96(a[i], b[j], c0, c1, c2);
        }
        if (k < N - 1)
            st |= c0;
        else
            H[k - (N - 1)] = c0;
        c0 = c1;
        c1 = c2;
        c2 = 0;
    }
    H[N] = c0;
    S = (uint32_t)(st != 0);
}

/* Square: off-diagonal products once, doubled, plus the diagonal. */
template <int N>
MP_HD void mant_sqr(const uint32_t* a, uint32_t* H, uint32_t& S)
{
    uint32_t c0 = 0, c1 = 0, c2 = 0, st = 0;
#pragma unroll
    for (int k = 0; k < 2 * N - 1; ++k) {
        uint32_t o0 = 0, o1 = 0, o2 = 0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            const int j = k - i;
            if (j > i && j < N)
                mac96(a[i], a[j], o0, o1, o2);
        }
        /* acc += 2*off */
        uint32_t d2 = (o2 << 1) | (o1 >> 31);
        uint32_t d1 = (o1 << 1) | (o0 >> 31);
        uint32_t d0 = o0 << 1;
        uint64_t t0 = (uint64_t)c0 + d0;
        uint64_t t1 = (uint64_t)c1 + d1 + (t0 >> 32);
        c0 = (uint32_t)t0;
        c1 = (uint32_t)t1;
        c2 = c2 + d2 + (uint32_t)(t1 >> 32);
        if ((k & 1) == 0)
            mac96(a[k >> 1], a[k >> 1], c0, c1, c2);
        if (k < N - 1)
            st |= c0;
        else
            H[k - (N - 1)] = c0;
        c0 = c1;
        c1 = c2;
        c2 = 0;
    }
    H[N] = c0;
    S = (uint32_t)(st != 0);
}

/* r = RN_p(a * b) (SQR: b ignored, a*a) */
template <int N, bool SQR>
MP_HD void mul_rn(mp<N>& r, const mp<N>& a, const mp<N>& b, int kclr)
{
    uint32_t H[N + 1];
    uint32_t S;
    if (SQR)
        mant_sqr<N>(a.m, H, S);
    else
        mant_mul<N>(a.m, b.m, H, S);
    int32_t e = SQR ? (a.e + a.e) : (a.e + b.e);
    const uint32_t s = SQR ? 0u : (a.s ^ b.s);
    /* product of normalised values is >= 2^(64N-2): at most one left shift */
    const bool sh = (H[N] >> 31) == 0;
#pragma unroll
    for (int j = N; j >= 1; --j)
        H[j] = sh ? ((H[j] << 1) | (H[j - 1] >> 31)) : H[j];
    H[0] = sh ? (H[0] << 1) : H[0];
    e -= sh ? 1 : 0;
    round_p<N>(H + 1, H[0], S, e, kclr);
    const bool z = (H[N] == 0);
#pragma unroll
    for (int j = 0; j < N; ++j) r.m[j] = H[j + 1];
    r.e = z ? EZERO : e;
    r.s = z ? 0u : s;
}

/* ------------------------------------------------------------------ */
/* r = RN_p(a + (-1)^flipb * b)                                         */
/* ------------------------------------------------------------------ */
template <int N>
MP_HD void add_rn(mp<N>& r, const mp<N>& a_in, const mp<N>& b_in, uint32_t flipb, int kclr)
{
    /* order by exponent: a gets the larger one */
    const bool sw = b_in.e > a_in.e;
    uint32_t A[N], B[N];
#pragma unroll
    for (int j = 0; j < N; ++j) {
        A[j] = sw ? b_in.m[j] : a_in.m[j];
        B[j] = sw ? a_in.m[j] : b_in.m[j];
    }
    const uint32_t sbin = b_in.s ^ flipb;
    const uint32_t sa = sw ? sbin : a_in.s;
    const uint32_t sb = sw ? a_in.s : sbin;
    const int32_t ea = sw ? b_in.e : a_in.e;
    const int32_t eb = sw ? a_in.e : b_in.e;

    /* align B right by d bits into N+1 words X (X[0] = guard word) + sticky */
    uint32_t d = (uint32_t)(ea - eb);
    if (d > 32u * (N + 1)) d = 32u * (N + 1);
    const uint32_t ws = d >> 5, bs = d & 31u;
    uint32_t X[N + 1];
    X[0] = funnel_r(B[0], 0u, bs);
#pragma unroll
    for (int j = 1; j < N; ++j) X[j] = funnel_r(B[j], B[j - 1], bs);
    X[N] = B[N - 1] >> bs;
    uint32_t S = 0;
#pragma unroll
    for (int st = 0; st < stages(N + 1); ++st) {
        const int t = 1 << st;
        const bool c = (ws & (uint32_t)t) != 0;
        if (wave_any(c)) {
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
    }
    S = (uint32_t)(S != 0);

    /* T = A*2^32 +/- (X + S): one carry chain, X complemented for a
       subtraction with carry-in 1 - S (A - X - S = A + ~X + 1 - S). */
    const uint32_t sub = sa ^ sb;
    const uint32_t msk = 0u - sub;
    uint32_t T[N + 1];
    uint32_t c = sub & (S ^ 1u);
    {
        uint64_t t = (uint64_t)(X[0] ^ msk) + c;
        T[0] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
    }
#pragma unroll
    for (int j = 1; j <= N; ++j) {
        uint64_t t = (uint64_t)A[j - 1] + (X[j] ^ msk) + c;
        T[j] = (uint32_t)t;
        c = (uint32_t)(t >> 32);
    }
    int32_t e = ea;
    uint32_t s = sa;

    /* add with carry-out: shift right one bit */
    const bool rs = (sub == 0) & (c != 0);
    S |= rs ? (T[0] & 1u) : 0u;
#pragma unroll
    for (int j = 0; j < N; ++j) T[j] = rs ? funnel_r(T[j + 1], T[j], 1u) : T[j];
    T[N] = rs ? ((T[N] >> 1) | 0x80000000u) : T[N];
    e += rs ? 1 : 0;

    /* subtract with borrow (only when d == 0, so S == 0): negate */
    const bool ng = (sub != 0) & (c == 0);
    if (wave_any(ng)) {
        uint32_t cc = ng ? 1u : 0u;
        const uint32_t nm = ng ? 0xffffffffu : 0u;
#pragma unroll
        for (int j = 0; j <= N; ++j) {
            uint64_t t = (uint64_t)(T[j] ^ nm) + cc;
            T[j] = (uint32_t)t;
            cc = (uint32_t)(t >> 32);
        }
        s = ng ? (s ^ 1u) : s;
    }

    /* normalise: 0 or 1 bit of left shift is the common case */
    const uint32_t top = T[N];
    const bool l1 = (top >> 30) == 1u;
    const bool big = (top >> 30) == 0u;   /* zero, or cancellation of >= 2 bits */
#pragma unroll
    for (int j = N; j >= 1; --j) T[j] = l1 ? funnel_r(T[j], T[j - 1], 31u) : T[j];
    T[0] = l1 ? (T[0] << 1) : T[0];
    e -= l1 ? 1 : 0;

    bool zero = false;
    if (wave_any(big)) {
        /* exact (S == 0) result with >= 2 leading zero bits, or zero */
        uint32_t nz = 0, lzw = 0;
        bool found = false;
#pragma unroll
        for (int j = N; j >= 0; --j) {
            const bool here = !found && (T[j] != 0);
            lzw = here ? (uint32_t)(N - j) : lzw;
            found = found || (T[j] != 0);
        }
        nz = found ? 1u : 0u;
        zero = big && !nz;
        const uint32_t wsh = big ? lzw : 0u;
#pragma unroll
        for (int st = 0; st < stages(N); ++st) {
            const int t = 1 << st;
            const bool cnd = (wsh & (uint32_t)t) != 0;
            if (wave_any(cnd)) {
#pragma unroll
                for (int j = N; j >= 0; --j) {
                    const uint32_t nv = (j - t >= 0) ? T[j - t] : 0u;
                    T[j] = cnd ? nv : T[j];
                }
            }
        }
        const uint32_t bsh = big ? clz32(T[N]) & 31u : 0u;
        if (wave_any(bsh != 0)) {
#pragma unroll
            for (int j = N; j >= 1; --j)
                T[j] = bsh ? funnel_r(T[j], T[j - 1], 32u - bsh) : T[j];
            T[0] = bsh ? (T[0] << bsh) : T[0];
        }
        e -= big ? (int32_t)(32u * wsh + bsh) : 0;
    }

    round_p<N>(T + 1, T[0], S, e, kclr);
#pragma unroll
    for (int j = 0; j < N; ++j) r.m[j] = zero ? 0u : T[j + 1];
    r.e = zero ? EZERO : e;
    r.s = zero ? 0u : s;
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

} /* namespace mdzg */

#endif
