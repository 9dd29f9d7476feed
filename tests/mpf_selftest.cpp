// Host build of mdz_amd/csrc/mpf_gmp.h checked against libgmp's mpf layer.
//
// Every operation of the GMP escape loop (src/frac_mandel.c:55-80 and
// siblings) -- mpf_add and mpf_sub (each alone and both through the one
// carry chain of gf_addsub), mpf_mul (distinct operands and squares),
// mpf_mul_ui by 2, mpf_abs, mpf_cmp with 4 and the escape test's shortcut on
// the integer limbs -- is compared limb for limb
// (size, exponent, sign, limbs) with GMP on operands written straight into
// mpf_t structures: random limbs, all-ones and zero limbs, equal leading
// limbs, the x+1/x and 1 00../0 ff.. near-cancellation patterns, every
// exponent distance up to past the precision, and every size up to PREC+1.
// The escape loops of the four formulas are compared step by step too.
//
// usage: mpf_selftest <seed> <cases-per-precision>
#include <gmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../mdz_amd/csrc/mpf_gmp.h"

using namespace mdzg::gmpf;

static uint64_t rng_state;
static uint64_t rnd64()
{
    uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static long fails = 0;

template <int L>
static void to_gmp(mpf_t x, const gf<L>& a)
{
    for (int i = 0; i < a.size; ++i) x->_mp_d[a.size - 1 - i] = a.t[i];
    x->_mp_size = a.neg ? -a.size : a.size;
    x->_mp_exp = a.size ? a.exp : 0;
}

template <int L>
static void from_gmp(gf<L>& a, const mpf_t x)
{
    const int n = x->_mp_size < 0 ? -x->_mp_size : x->_mp_size;
    for (int i = 0; i < L; ++i) a.t[i] = (i < n) ? x->_mp_d[n - 1 - i] : 0;
    a.size = n;
    a.exp = n ? (int32_t)x->_mp_exp : 0;
    a.neg = x->_mp_size < 0;
}

// equal values, each with zero limbs beyond its size (the representation
// the device's full-width forms rely on)
template <int L>
static bool same(const gf<L>& a, const gf<L>& b)
{
    for (int i = 0; i < L; ++i)
        if ((i >= a.size && a.t[i] != 0) || (i >= b.size && b.t[i] != 0)) return false;
    if (a.size != b.size) return false;
    if (a.size == 0) return true;
    if (a.exp != b.exp || a.neg != b.neg) return false;
    for (int i = 0; i < a.size; ++i)
        if (a.t[i] != b.t[i]) return false;
    return true;
}

static uint64_t rnd_limb()
{
    switch (rnd64() % 8) {
    case 0: return 0;
    case 1: return ~0ull;
    case 2: return 1;
    case 3: return 1ull << (rnd64() % 64);
    default: return rnd64();
    }
}

template <int L>
static void rand_op(gf<L>& a, int emin, int emax, int lr = L)
{
    a.size = 1 + (int)(rnd64() % (uint64_t)lr);
    for (int i = 0; i < L; ++i) a.t[i] = (i < a.size) ? rnd_limb() : 0;
    if (a.t[0] == 0) a.t[0] = 1 + (rnd64() % 7);
    a.exp = emin + (int)(rnd64() % (uint64_t)(emax - emin + 1));
    a.neg = rnd64() & 1;
    if (rnd64() % 32 == 0) { for (int i = 0; i < L; ++i) a.t[i] = 0; a.size = 0; a.exp = 0; a.neg = 0; }
}

// b close to a: same leading limbs, then a near-cancellation pattern
template <int L>
static void near_op(gf<L>& b, const gf<L>& a, int lr = L)
{
    b = a;
    if (a.size == 0) { rand_op(b, -3, 3, lr); return; }
    const int k = (int)(rnd64() % (uint64_t)(a.size + 1));   // equal leading limbs
    b.size = k + (int)(rnd64() % (uint64_t)(lr - k + 1));
    if (b.size == 0) b.size = 1;
    for (int i = k; i < L; ++i) b.t[i] = (i < b.size) ? rnd_limb() : 0;
    switch (rnd64() % 4) {
    case 0:   // x+1 / x then zeros vs all-ones
        if (k < b.size && k < a.size && a.t[k] > 0) {
            b.t[k] = a.t[k] - 1;
            for (int i = k + 1; i < b.size; ++i) b.t[i] = (rnd64() % 4) ? ~0ull : rnd64();
        }
        break;
    case 1:   // 1 00.. (a, one limb higher) vs 0 ff.. (b)
        if (k == 0) {
            b.exp = a.exp - 1;
            b.t[0] = ~0ull;
            for (int i = 1; i < b.size; ++i) b.t[i] = (rnd64() % 4) ? ~0ull : rnd64();
        }
        break;
    default:
        break;
    }
    if (b.t[0] == 0) b.t[0] = 1;
    b.neg = rnd64() & 1;
}

template <int L>
static void report(const char* what, int P, const gf<L>& a, const gf<L>& b, const gf<L>& mine, const gf<L>& ref)
{
    if (fails++ >= 12) return;
    auto dump = [](const char* n, const gf<L>& x) {
        fprintf(stderr, "  %s size=%d exp=%d neg=%u :", n, x.size, x.exp, x.neg);
        for (int i = 0; i < x.size; ++i) fprintf(stderr, " %016llx", (unsigned long long)x.t[i]);
        fprintf(stderr, "\n");
    };
    fprintf(stderr, "MISMATCH %s PREC=%d\n", what, P);
    dump("a   ", a);
    dump("b   ", b);
    dump("mine", mine);
    dump("gmp ", ref);
}

// LR = PREC + 1 limbs (the semantics); values stored in gf<S>, S >= LR
template <int LR, int S = LR>
static void check(long cases)
{
    constexpr int L = S;
    constexpr int P = LR - 1;
    const unsigned long bits = 64ul * P - 127ul > 53 ? 64ul * P - 127ul : 53;
    mpf_t A, B, R;
    mpf_init2(A, bits);
    mpf_init2(B, bits);
    mpf_init2(R, bits);
    if (A->_mp_prec != P) {
        fprintf(stderr, "unexpected PREC %d for %lu bits (want %d)\n", A->_mp_prec, bits, P);
        ++fails;
        return;
    }
    for (long it = 0; it < cases; ++it) {
        gf<L> a, b, r, ref;
        rand_op(a, -4, 4, LR);
        if (rnd64() % 2) near_op(b, a, LR);
        else {
            const int ea = a.size ? a.exp : 0;
            rand_op(b, ea - (LR + 2), ea + (LR + 2), LR);
        }
        to_gmp(A, a);
        to_gmp(B, b);
        gf_add<L>(r, a, b, 0u, LR); mpf_add(R, A, B); from_gmp(ref, R);
        if (!same(r, ref)) report("add", P, a, b, r, ref);
        gf_sub<L>(r, a, b, LR); mpf_sub(R, A, B); from_gmp(ref, R);
        if (!same(r, ref)) report("sub", P, a, b, r, ref);
        gf_addsub<L>(r, a, b, 1u, LR);
        if (!same(r, ref)) report("addsub-sub", P, a, b, r, ref);
        mpf_add(R, A, B); from_gmp(ref, R);
        gf_addsub<L>(r, a, b, 0u, LR);
        if (!same(r, ref)) report("addsub-add", P, a, b, r, ref);
        {
            // the escape test's shortcut on nonnegative operands
            gf<L> x = a, y = b;
            x.neg = 0;
            y.neg = 0;
            if (rnd64() % 2) {        // integer parts around the undecided 3..4
                x.exp = 1; x.t[0] = rnd64() % 6 + (x.t[0] == 0);
                if (x.size == 0) { x.size = 1; }
                y.exp = (rnd64() % 3) ? 1 : (int)(rnd64() % 3) - 1;
                if (y.size == 0) { y.size = 1; }
                y.t[0] = (y.exp == 1) ? rnd64() % 6 + 1 : (y.t[0] | 1);
            }
            mpf_t X, Y, F;
            mpf_init2(X, bits);
            mpf_init2(Y, bits);
            mpf_init2(F, bits);
            mpf_set_si(F, 4);
            to_gmp(X, x);
            to_gmp(Y, y);
            mpf_add(R, X, Y);
            const int want = mpf_cmp(R, F) > 0;
            const int q = gf_sum_gt4_quick<L>(x, y);
            if (q >= 0 && q != want) report("sum-gt4-quick", P, x, y, x, y);
            mpf_clear(X);
            mpf_clear(Y);
            mpf_clear(F);
        }
        gf_mul<L>(r, a, b, LR); mpf_mul(R, A, B); from_gmp(ref, R);
        if (!same(r, ref)) report("mul", P, a, b, r, ref);
        gf_mul<L>(r, a, a, LR); mpf_mul(R, A, A); from_gmp(ref, R);
        if (!same(r, ref)) report("sqr", P, a, a, r, ref);
        gf_sqr<L>(r, a, LR);
        if (!same(r, ref)) report("sqr-sym", P, a, a, r, ref);
        gf_mul2<L>(r, a, LR); mpf_mul_ui(R, A, 2); from_gmp(ref, R);
        if (!same(r, ref)) report("mul2", P, a, a, r, ref);
        {
            mpf_t F;
            mpf_init2(F, bits);
            mpf_set_si(F, 4);
            gf<L> x = a;
            x.neg = 0;
            to_gmp(R, x);
            const bool mine = gf_gt4<L>(x), want = mpf_cmp(R, F) > 0;
            if (mine != want) report("gt4", P, x, x, x, x);
            mpf_clear(F);
        }
    }
    mpf_clear(A);
    mpf_clear(B);
    mpf_clear(R);
}

// the GMP escape loops (src/frac_mandel.c:55-80 and siblings), step by step
template <int L>
static void check_chain(long cases)
{
    constexpr int P = L - 1;
    const unsigned long bits = 64ul * P - 127ul > 53 ? 64ul * P - 127ul : 53;
    mpf_t wre, wim, wre2, wim2, cre, cim, t1, bail;
    mpf_t* all[] = {&wre, &wim, &wre2, &wim2, &cre, &cim, &t1, &bail};
    for (mpf_t* x : all) mpf_init2(*x, bits);
    mpf_set_si(bail, 4);
    for (long it = 0; it < cases; ++it) {
        const int fractal = (int)(rnd64() % 4);
        gf<L> cr, ci;
        rand_op(cr, 0, 1);
        rand_op(ci, -1, 1);
        cr.size = L; ci.size = L;
        for (int i = 0; i < L; ++i) { cr.t[i] = rnd64(); ci.t[i] = rnd64(); }
        cr.t[0] = (cr.exp == 1) ? 1 : (cr.t[0] | 1);
        ci.t[0] = (ci.exp == 1) ? 1 : (ci.t[0] | 1);
        to_gmp(cre, cr);
        to_gmp(cim, ci);
        mpf_set(wre, cre);
        mpf_set(wim, cim);
        mpf_mul(wre2, wre, wre);
        mpf_mul(wim2, wim, wim);
        gf<L> zr, zi, zr2, zi2, t, s;
        from_gmp(zr, wre); from_gmp(zi, wim); from_gmp(zr2, wre2); from_gmp(zi2, wim2);
        for (int wz = 1; wz <= 300; ++wz) {
            mpf_mul(t1, wre, wim);
            if (fractal == 1) mpf_abs(t1, t1);
            mpf_mul_ui(t1, t1, 2);
            mpf_add(wim, t1, cim);
            mpf_sub(t1, wre2, wim2);
            if (fractal == 2 || (fractal == 3 && (wz % 2))) mpf_abs(t1, t1);
            mpf_add(wre, t1, cre);
            mpf_mul(wim2, wim, wim);
            mpf_mul(wre2, wre, wre);
            mpf_add(t1, wim2, wre2);
            const int esc_ref = mpf_cmp(t1, bail) > 0;

            gf_mul<L>(t, zr, zi, L);
            if (fractal == 1) gf_abs<L>(t);
            gf<L> t2;
            gf_mul2<L>(t2, t, L);
            gf_addsub<L>(zi, t2, ci, 0u, L);
            gf_addsub<L>(t, zr2, zi2, 1u, L);
            if (fractal == 2 || (fractal == 3 && (wz % 2))) gf_abs<L>(t);
            gf_addsub<L>(zr, t, cr, 0u, L);
            gf_sqr<L>(zi2, zi, L);
            gf_sqr<L>(zr2, zr, L);
            int esc = gf_sum_gt4_quick<L>(zi2, zr2);
            if (esc < 0) {
                gf_add<L>(s, zi2, zr2, 0u, L);
                esc = gf_gt4<L>(s) ? 1 : 0;
            }

            gf<L> rr, ri;
            from_gmp(rr, wre);
            from_gmp(ri, wim);
            if (!same(rr, zr) || !same(ri, zi) || esc != esc_ref) {
                if (fails < 12)
                    fprintf(stderr, "CHAIN MISMATCH PREC=%d fractal=%d step=%d\n", P, fractal, wz);
                ++fails;
                break;
            }
            if (esc_ref) break;
        }
    }
    for (mpf_t* x : all) mpf_clear(*x);
}

int main(int argc, char** argv)
{
    rng_state = argc > 1 ? strtoull(argv[1], 0, 0) : 1;
    long cases = argc > 2 ? atol(argv[2]) : 20000;
    check<2>(cases);
    check<3>(cases);
    check<4>(cases);
    check<5>(cases);
    check<6>(cases);
    check<7>(cases);
    check<9>(cases);
    check<18>(cases / 4);
    check<3, 4>(cases / 2);    // the device's storage buckets hold fewer limbs in use
    check<4, 6>(cases / 2);
    check<6, 10>(cases / 2);
    check<9, 18>(cases / 4);
    check_chain<3>(cases / 20);
    check_chain<4>(cases / 20);
    check_chain<6>(cases / 20);
    check_chain<18>(cases / 80);
    printf("%s fails=%ld\n", fails ? "FAIL" : "OK", fails);
    return fails ? 1 : 0;
}
