// Host check of mdz_amd/csrc/colour_math.h, the colour kernel's integer
// helpers: the multiply-based remainder against %, and (guint32)v against its
// definition (the low 32 bits of the 64-bit truncation) on both sides of 2^32.
// Prints "ok" or the first mismatch.
#include <cstdio>
#include <cstdlib>
#include <random>

#include "../mdz_amd/csrc/colour_math.h"

using namespace mdzg::colour;

int main(int argc, char** argv)
{
    const int n = argc > 1 ? atoi(argv[1]) : 100000;
    std::mt19937_64 rng(12345);
    const uint32_t ds[] = { 1, 2, 3, 5, 7, 16, 36, 99, 127, 128, 199, 254, 255, 256, 65535,
                            1000003, 0x7fffffffu, 0xfffffffeu, 0xffffffffu };
    for (uint32_t d : ds) {
        const uint64_t M = rem_magic(d);
        const uint32_t edge[] = { 0, 1, d - 1, d, d + 1, 2 * d, 0xffffffffu, 0xfffffffeu,
                                  0x80000000u, 0x7fffffffu };
        for (uint32_t a : edge)
            if (rem(a, M, d) != a % d) {
                printf("rem %u %% %u: %u\n", a, d, rem(a, M, d));
                return 1;
            }
        for (int i = 0; i < n; ++i) {
            const uint32_t a = (uint32_t)rng();
            if (rem(a, M, d) != a % d) {
                printf("rem %u %% %u: %u\n", a, d, rem(a, M, d));
                return 1;
            }
        }
    }
    for (uint32_t d = 1; d <= 256; ++d) {
        const uint64_t M = rem_magic(d);
        for (uint32_t a = 0; a < 70000; ++a)
            if (rem(a, M, d) != a % d) {
                printf("rem %u %% %u\n", a, d);
                return 1;
            }
    }
    std::uniform_real_distribution<double> u01(0.0, 1.0);
    for (int i = 0; i < n; ++i) {
        const double scale = ldexp(1.0, (int)(rng() % 70));            // up to 2^69
        double v = floor(u01(rng) * scale) + (rng() % 4) * 0.25;
        if (rng() % 3 == 0)
            v = -v;
        const uint32_t want = (v > -9223372036854775808.0 && v < 9223372036854775808.0)
                                  ? (uint32_t)(uint64_t)(int64_t)v : 0u;
        if (u32_of(v) != want) {
            printf("u32_of(%.17g) = %u, want %u\n", v, u32_of(v), want);
            return 1;
        }
    }
    const double edges[] = { 0.0, 0.5, 1.0, 4294967295.0, 4294967295.5, 4294967296.0,
                             4294967297.0, 9007199254740993.0, 9223372036854774784.0,
                             9223372036854775808.0, 1e300, -0.5, -1.0, -4294967296.0,
                             -9223372036854775808.0, -1e300 };
    for (double v : edges) {
        const uint32_t want = (v > -9223372036854775808.0 && v < 9223372036854775808.0)
                                  ? (uint32_t)(uint64_t)(int64_t)v : 0u;
        if (u32_of(v) != want) {
            printf("u32_of(%.17g) = %u, want %u\n", v, u32_of(v), want);
            return 1;
        }
    }
    printf("ok\n");
    return 0;
}
