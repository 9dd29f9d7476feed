// Host build of mdz_amd/csrc/mp_rn.h checked against MPFR (MPFR_RNDN).
//
// Every add, sub, mul and sqr of the header is compared bit for bit with
// mpfr_add / mpfr_sub / mpfr_mul at the same precision, over random and
// edge-case operands: all alignment distances up to past the guard word,
// massive cancellation, ties (round-half-even), all-ones significands that
// carry out on rounding, zeros, every precision from 1 to 32N for small N
// and the word-aligned and ragged precisions for large N.  A short formula
// chain (the escape loop of src/frac_mandel.c:33-51) is compared too.
//
// usage: mp_selftest <seed> <cases-per-config>
#include <gmp.h>
#include <mpfr.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MDZG_COUNT_SHORT_FALLBACK 1
#include "../mdz_amd/csrc/mp_rn.h"

long mdzg::mdzg_short_fallbacks = 0;

using namespace mdzg;

static uint64_t rng_state;
static uint64_t rnd64()
{
    // splitmix64
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static uint32_t rnd32() { return (uint32_t)rnd64(); }

template <int N>
static void to_mpfr(mpfr_t out, const mp<N>& a)
{
    mpfr_set_prec(out, 32 * N);
    if (is_zero(a)) {
        mpfr_set_zero(out, a.s ? -1 : 1);
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpz_import(z, N, -1, 4, 0, 0, a.m);
    mpfr_set_z_2exp(out, z, (mpfr_exp_t)a.e - 32 * N, MPFR_RNDN);
    if (a.s) mpfr_neg(out, out, MPFR_RNDN);
    mpz_clear(z);
}

template <int N>
static void from_mpfr(mp<N>& r, mpfr_t x)
{
    memset(r.m, 0, sizeof r.m);
    if (mpfr_zero_p(x)) {
        r.e = EZERO;
        r.s = 0;
        return;
    }
    mpz_t z;
    mpz_init(z);
    mpfr_exp_t ex = mpfr_get_z_2exp(z, x);
    mpz_abs(z, z);
    long prec = (long)mpfr_get_prec(x);
    mpz_mul_2exp(z, z, 32 * N - prec);
    size_t cnt = 0;
    mpz_export(r.m, &cnt, -1, 4, 0, 0, z);
    r.e = (int32_t)(ex + prec);
    r.s = mpfr_signbit(x) ? 1u : 0u;
    mpz_clear(z);
}

template <int N>
static bool same(const mp<N>& a, const mp<N>& b)
{
    if (is_zero(a) && is_zero(b)) return true;   // sign of zero is not compared
    if (a.e != b.e || a.s != b.s) return false;
    for (int j = 0; j < N; ++j)
        if (a.m[j] != b.m[j]) return false;
    return true;
}

// random p-bit operand
template <int N>
static void rand_op(mp<N>& a, int p, int emin, int emax)
{
    int kind = (int)(rnd32() % 16);
    for (int j = 0; j < N; ++j) a.m[j] = rnd32();
    if (kind == 0) {            // power of two
        for (int j = 0; j < N; ++j) a.m[j] = 0;
    } else if (kind == 1) {     // all ones
        for (int j = 0; j < N; ++j) a.m[j] = 0xffffffffu;
    } else if (kind == 2) {     // sparse: few bits
        for (int j = 0; j < N; ++j) a.m[j] = (rnd32() % 4 == 0) ? (1u << (rnd32() % 32)) : 0u;
    } else if (kind == 4 || kind == 5) {
        // top and bottom words only; kind 5 with top word 2^31: products of
        // two such values put 0 or 2^31 in the guard word (short-product
        // checks of mul_rn fall back to the full product there)
        for (int j = 1; j < N - 1; ++j) a.m[j] = 0;
        if (kind == 5) a.m[N - 1] = 0x80000000u;
    }
    a.m[N - 1] |= 0x80000000u;
    int kclr = 32 * N - p;
    for (int b = 0; b < kclr; ++b) a.m[b >> 5] &= ~(1u << (b & 31));
    if (kind == 6 && kclr % 2 == 0) {
        // 2^(32N-1) + 2^(16N + kclr/2 - 1): its square is an exact tie at p
        // with an even last bit
        for (int j = 0; j < N; ++j) a.m[j] = 0;
        const int t = 16 * N + kclr / 2 - 1;
        a.m[N - 1] = 0x80000000u;
        a.m[t >> 5] |= 1u << (t & 31);
    }
    a.e = emin + (int)(rnd32() % (uint32_t)(emax - emin + 1));
    a.s = rnd32() & 1u;
    if (kind == 3) set_zero(a);
}

// b close to a (cancellation): a with a few low bits changed, same exponent or off by one
template <int N>
static void near_op(mp<N>& b, const mp<N>& a, int p)
{
    if (is_zero(a)) {
        rand_op(b, p, -6, 6);
        return;
    }
    b = a;
    if (rnd32() % 8 == 0) {
        // exact copy: a - a cancels to zero, a + a doubles
        b.s = rnd32() & 1u;
        return;
    }
    int kclr = 32 * N - p;
    int nb = 1 + (int)(rnd32() % (uint32_t)(p < 70 ? p : 70));
    for (int i = 0; i < 3; ++i) {
        int bit = kclr + (int)(rnd32() % (uint32_t)nb);
        if (bit < 32 * N - 1) b.m[bit >> 5] ^= (1u << (bit & 31));
    }
    if (rnd32() % 4 == 0) {
        // b = a - ulp(a) pattern across a binade: mantissa all ones, exponent - 1
        for (int j = 0; j < N; ++j) b.m[j] = 0xffffffffu;
        for (int bb = 0; bb < kclr; ++bb) b.m[bb >> 5] &= ~(1u << (bb & 31));
        b.e = a.e - 1;
    }
    b.s = rnd32() & 1u;
}

static long fails = 0;

template <int N>
static void check_p(int p, long cases)
{
    const int kclr = 32 * N - p;
    mpfr_t A, B, R;
    mpfr_inits2(32 * N, A, B, (mpfr_ptr)0);
    mpfr_init2(R, p);
    for (long it = 0; it < cases; ++it) {
        mp<N> a, b, r, ref;
        rand_op(a, p, -6, 6);
        int mode = (int)(rnd32() % 4);
        if (mode == 0) {
            near_op(b, a, p);
        } else {
            int spread = (mode == 1) ? 4 : (mode == 2 ? 32 * (N + 3) : 3000);
            int ec = is_zero(a) ? 0 : a.e;
            rand_op(b, p, ec - spread, ec + spread);
        }
        to_mpfr(A, a);
        to_mpfr(B, b);
        for (int op = 0; op < 5; ++op) {
            if (op == 0) { add_rn<N>(r, a, b, 0u, kclr); mpfr_add(R, A, B, MPFR_RNDN); }
            else if (op == 1) { add_rn<N>(r, a, b, 1u, kclr); mpfr_sub(R, A, B, MPFR_RNDN); }
            else if (op == 2) { mul_rn<N, false>(r, a, b, kclr); mpfr_mul(R, A, B, MPFR_RNDN); }
            else if (op == 3) { mul_rn<N, true>(r, a, a, kclr); mpfr_mul(R, A, A, MPFR_RNDN); }
            else {
                /* |a| - |b| through the specialisation for nonnegative operands */
                mp<N> pa = a, pb = b;
                pa.s = 0;
                pb.s = 0;
                add_rn<N, SUB_NONNEG>(r, pa, pb, 1u, kclr);
                if (kclr == 0) {   // the word-aligned kernels' form agrees
                    mp<N> r2;
                    add_rn<N, SUB_NONNEG, true>(r2, pa, pb, 1u, kclr);
                    if (!same(r2, r) && fails++ < 10)
                        fprintf(stderr, "SUB_NONNEG WORD-ALIGNED MISMATCH N=%d\n", N);
                }
                mpfr_t PA, PB;
                mpfr_inits2(32 * N, PA, PB, (mpfr_ptr)0);
                mpfr_abs(PA, A, MPFR_RNDN);
                mpfr_abs(PB, B, MPFR_RNDN);
                mpfr_sub(R, PA, PB, MPFR_RNDN);
                mpfr_clears(PA, PB, (mpfr_ptr)0);
            }
            from_mpfr(ref, R);
            if (!same(r, ref)) {
                if (fails < 10) {
                    fprintf(stderr, "MISMATCH N=%d p=%d op=%d\n a=", N, p, op);
                    mpfr_out_str(stderr, 16, 0, A, MPFR_RNDN);
                    fprintf(stderr, "\n b=");
                    mpfr_out_str(stderr, 16, 0, B, MPFR_RNDN);
                    fprintf(stderr, "\n mpfr=");
                    mpfr_out_str(stderr, 16, 0, R, MPFR_RNDN);
                    mpfr_t G;
                    mpfr_init2(G, 32 * N);
                    to_mpfr(G, r);
                    fprintf(stderr, "\n mine=");
                    mpfr_out_str(stderr, 16, 0, G, MPFR_RNDN);
                    fprintf(stderr, "\n");
                    mpfr_clear(G);
                }
                ++fails;
            }
        }
    }
    mpfr_clears(A, B, R, (mpfr_ptr)0);
}

// escape-loop chain (Mandelbrot ordering) at precision p, compared per step
template <int N>
static void check_chain(int p, long cases)
{
    const int kclr = 32 * N - p;
    mpfr_t wre, wim, wre2, wim2, cre, cim, t1, tmp;
    mpfr_inits2(p, wre, wim, wre2, wim2, cre, cim, t1, (mpfr_ptr)0);
    mpfr_init2(tmp, 32 * N);
    for (long it = 0; it < cases; ++it) {
        mp<N> x, y;
        rand_op(x, p, -3, 1);
        rand_op(y, p, -40, 1);
        to_mpfr(tmp, x); mpfr_set(cre, tmp, MPFR_RNDN);
        to_mpfr(tmp, y); mpfr_set(cim, tmp, MPFR_RNDN);
        mpfr_set(wre, cre, MPFR_RNDN);
        mpfr_set(wim, cim, MPFR_RNDN);
        mpfr_mul(wre2, wre, wre, MPFR_RNDN);
        mpfr_mul(wim2, wim, wim, MPFR_RNDN);
        mp<N> zr = x, zi = y, cr = x, ci = y, zr2, zi2, t, s;
        mul_rn<N, true>(zr2, zr, zr, kclr);
        mul_rn<N, true>(zi2, zi, zi, kclr);
        for (int wz = 1; wz <= 200; ++wz) {
            mpfr_mul(t1, wre, wim, MPFR_RNDN);
            mpfr_mul_si(t1, t1, 2, MPFR_RNDN);
            mpfr_add(wim, t1, cim, MPFR_RNDN);
            mpfr_sub(t1, wre2, wim2, MPFR_RNDN);
            mpfr_add(wre, t1, cre, MPFR_RNDN);
            mpfr_mul(wim2, wim, wim, MPFR_RNDN);
            mpfr_mul(wre2, wre, wre, MPFR_RNDN);
            mpfr_add(t1, wim2, wre2, MPFR_RNDN);
            int esc_ref = mpfr_cmp_si(t1, 4) > 0;

            mul_rn<N, false>(t, zr, zi, kclr);
            twice(t);
            add_rn<N>(zi, t, ci, 0u, kclr);
            if (kclr == 0)
                add_rn<N, SUB_NONNEG, true>(t, zr2, zi2, 1u, kclr);
            else
                add_rn<N, SUB_NONNEG>(t, zr2, zi2, 1u, kclr);
            add_rn<N>(zr, t, cr, 0u, kclr);
            mul_rn<N, true>(zi2, zi, zi, kclr);
            mul_rn<N, true>(zr2, zr, zr, kclr);
            add_rn<N>(s, zi2, zr2, 0u, kclr);
            int esc = gt4(s) ? 1 : 0;

            mp<N> rr, ri;
            from_mpfr(rr, wre);
            from_mpfr(ri, wim);
            if (!same(rr, zr) || !same(ri, zi) || esc != esc_ref) {
                if (fails < 10)
                    fprintf(stderr, "CHAIN MISMATCH N=%d p=%d step=%d esc=%d/%d\n", N, p, wz, esc, esc_ref);
                ++fails;
                break;
            }
            if (esc_ref) break;
        }
    }
    mpfr_clears(wre, wim, wre2, wim2, cre, cim, t1, tmp, (mpfr_ptr)0);
}

// escape decision RN(a+b) > 4 from approx_nonneg + exact fallback, with sums
// placed on, just above and just below the threshold 4 + 2^(2-p)
template <int N>
static void check_escape(int p, long cases)
{
    const int kclr = 32 * N - p;
    const double thresh = 4.0 + ldexp(1.0, 2 - p);
    mpfr_t A, B, R;
    mpfr_inits2(32 * N, A, B, (mpfr_ptr)0);
    mpfr_init2(R, p);
    for (long it = 0; it < cases; ++it) {
        mp<N> a, b;
        rand_op(a, p, -3, 3);
        a.s = 0;
        if (is_zero(a)) rand_op(a, p, 1, 2), a.s = 0;
        int mode = (int)(rnd32() % 3);
        if (mode == 0) {
            rand_op(b, p, -3, 3);
        } else {
            // b = RN_p(target - a) with target 4 + 2^(2-p) + k * 2^(e) small offsets
            mpfr_t T;
            mpfr_init2(T, 4 * 32 * N + 64);
            to_mpfr(A, a);
            mpfr_set_ui(T, 4, MPFR_RNDN);
            mpfr_t q;
            mpfr_init2(q, 64);
            mpfr_set_ui_2exp(q, 1, 2 - p, MPFR_RNDN);
            if (mode == 2) mpfr_add(T, T, q, MPFR_RNDN);
            long k = (long)(rnd32() % 7) - 3;
            mpfr_set_si_2exp(q, k, -(long)(rnd32() % (uint32_t)(2 * p)), MPFR_RNDN);
            mpfr_add(T, T, q, MPFR_RNDN);
            mpfr_sub(T, T, A, MPFR_RNDN);
            if (mpfr_sgn(T) <= 0) { mpfr_clears(T, q, (mpfr_ptr)0); continue; }
            mpfr_t t2;
            mpfr_init2(t2, p);
            mpfr_set(t2, T, MPFR_RNDN);
            from_mpfr(b, t2);
            mpfr_clears(T, q, t2, (mpfr_ptr)0);
        }
        b.s = 0;
        to_mpfr(A, a);
        to_mpfr(B, b);
        mpfr_add(R, A, B, MPFR_RNDN);
        const bool ref = mpfr_cmp_ui(R, 4) > 0;
        const bool got = escaped<N>(a, b, thresh, kclr);
        if (ref != got) {
            if (fails < 10) fprintf(stderr, "ESCAPE MISMATCH N=%d p=%d mode=%d\n", N, p, mode);
            ++fails;
        }
    }
    mpfr_clears(A, B, R, (mpfr_ptr)0);
}

// the symmetric squares (off-diagonal products once, then one doubling pass
// with the diagonal squares; for the exact square also the per-column
// doubling form) against the plain product of a with itself: every output
// word, from column K0 up (K0 = 0, the full product, and the short
// products' first columns, odd ones included)
template <int N, int K0>
static void check_sqr_sym(long cases)
{
    for (long it = 0; it < cases; ++it) {
        uint32_t a[N];
        for (int j = 0; j < N; ++j) {
            const uint32_t r = rnd32();
            a[j] = (r % 5 == 0) ? 0u : (r % 5 == 1) ? ~0u : rnd32();
        }
        uint32_t H1[N + 1], H2[N + 1], H3[N + 1], S1, S2, S3;
        mant_sqr_sym<N, K0>(a, H1, S1);
        mant_mul<N, K0>(a, a, H2, S2);
        bool ok = S1 == S2;
        for (int j = 0; j <= N; ++j) ok = ok && H1[j] == H2[j];
        if (K0 == 0) {              // the exact fallback's per-column form
            mant_sqr_exact<N>(a, H3, S3);
            ok = ok && S3 == S2;
            for (int j = 0; j <= N; ++j) ok = ok && H3[j] == H2[j];
        }
        if (!ok) {
            if (fails < 10) fprintf(stderr, "SQR-SYM MISMATCH N=%d K0=%d\n", N, K0);
            ++fails;
        }
    }
}

template <int N>
static void check_n(long cases)
{
    int ps[8];
    int np = 0;
    ps[np++] = 32 * N;
    ps[np++] = 32 * N - 1;
    if (N > 1) ps[np++] = 32 * (N - 1) + 1;
    ps[np++] = 1 + (int)(rnd32() % (uint32_t)(32 * N));
    ps[np++] = 1 + (int)(rnd32() % (uint32_t)(32 * N));
    if (N <= 2)
        for (int p = 1; p <= 32 * N; ++p) check_p<N>(p, cases / 8 + 1);
    for (int i = 0; i < np; ++i) {
        check_escape<N>(ps[i], cases);
        check_p<N>(ps[i], cases);
        check_chain<N>(ps[i], cases / 20 + 1);
    }
}

int main(int argc, char** argv)
{
    rng_state = argc > 1 ? strtoull(argv[1], 0, 0) : 1;
    long cases = argc > 2 ? atol(argv[2]) : 2000;
    check_sqr_sym<1, 0>(cases);
    check_sqr_sym<2, 0>(cases);
    check_sqr_sym<2, 1>(cases);
    check_sqr_sym<3, 0>(cases);
    check_sqr_sym<5, 3>(cases);
    check_sqr_sym<12, 0>(cases);
    check_sqr_sym<12, 10>(cases);
    check_sqr_sym<16, 14>(cases);
    check_sqr_sym<24, 22>(cases);
    check_sqr_sym<32, 0>(cases);
    check_sqr_sym<32, 30>(cases);
    check_n<1>(cases);
    check_n<2>(cases);
    check_n<3>(cases);
    check_n<4>(cases);
    check_n<6>(cases);
    check_n<8>(cases);
    check_n<12>(cases);
    check_n<16>(cases / 2);
    check_n<24>(cases / 4);
    check_n<32>(cases / 4);
    // the precisions of the gallery scenes and benchmark classes
    check_p<3>(80, cases); check_chain<3>(80, cases / 10);
    check_p<6>(176, cases); check_p<6>(184, cases); check_chain<6>(184, cases / 10);
    check_p<10>(320, cases); check_chain<10>(320, cases / 10);
    printf("short-product fallbacks=%ld\n", mdzg::mdzg_short_fallbacks);
    printf("%s fails=%ld\n", fails ? "FAIL" : "OK", fails);
    return fails ? 1 : 0;
}
