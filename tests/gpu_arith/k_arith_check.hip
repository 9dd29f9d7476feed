// Device build of mdz_amd/csrc/mp_rn.h for the tests: one thread per operand
// pair runs the kernels' own add_rn / mul_rn, so that every lane mix of a
// wave (the votes, the short products' fallback, the inline-asm products) is
// checked op by op against MPFR on the host.
//
//   int arith_check(int nwords, int op, int kclr, int n,
//                   const uint32_t* a, const uint32_t* b, uint32_t* r)
//
// a, b, r: n stored values of MDZG_VS(nwords) words each (significand words,
// exponent, sign), host memory.  op: 0 a+b, 1 a-b, 2 a*b, 3 a*a, 4 a-b for
// nonnegative a and b (the SUB_NONNEG form).
// Returns 0, -1 for an unsupported word count or bad arguments, -2 on a HIP
// error.
#include <hip/hip_runtime.h>

#include "../../mdz_amd/csrc/mdzg_dev.h"
#include "../../mdz_amd/csrc/mp_rn.h"

using namespace mdzg;

template <int N>
__global__ __launch_bounds__(256) void k_check(int op, int kclr, int n, const uint32_t* a,
                                               const uint32_t* b, uint32_t* r)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    constexpr int VS = MDZG_VS(N);
    mp<N> x, y, z;
#pragma unroll
    for (int j = 0; j < N; ++j) {
        x.m[j] = a[(size_t)i * VS + j];
        y.m[j] = b[(size_t)i * VS + j];
    }
    x.e = (int32_t)a[(size_t)i * VS + N];
    x.s = a[(size_t)i * VS + N + 1];
    y.e = (int32_t)b[(size_t)i * VS + N];
    y.s = b[(size_t)i * VS + N + 1];
    if (op == 0)
        add_rn<N>(z, x, y, 0u, kclr);
    else if (op == 1)
        add_rn<N>(z, x, y, 1u, kclr);
    else if (op == 2)
        mul_rn<N, false>(z, x, y, kclr);
    else if (op == 3)
        mul_rn<N, true>(z, x, x, kclr);
    else if (kclr == 0)   /* as the word-aligned kernels call it */
        add_rn<N, SUB_NONNEG, true>(z, x, y, 1u, kclr);
    else
        add_rn<N, SUB_NONNEG>(z, x, y, 1u, kclr);
#pragma unroll
    for (int j = 0; j < N; ++j) r[(size_t)i * VS + j] = z.m[j];
    r[(size_t)i * VS + N] = (uint32_t)z.e;
    r[(size_t)i * VS + N + 1] = z.s;
}

template <int N>
static int run(int op, int kclr, int n, const uint32_t* a, const uint32_t* b, uint32_t* r)
{
    if (kclr < 0 || kclr >= 32 * N)
        return -1;
    const size_t bytes = (size_t)n * MDZG_VS(N) * 4;
    uint32_t *da = nullptr, *db = nullptr, *dr = nullptr;
    int rc = -2;
    if (hipMalloc(&da, bytes) == hipSuccess && hipMalloc(&db, bytes) == hipSuccess &&
        hipMalloc(&dr, bytes) == hipSuccess &&
        hipMemcpy(da, a, bytes, hipMemcpyHostToDevice) == hipSuccess &&
        hipMemcpy(db, b, bytes, hipMemcpyHostToDevice) == hipSuccess) {
        hipLaunchKernelGGL(k_check<N>, dim3((n + 255) / 256), dim3(256), 0, 0, op, kclr, n, da,
                           db, dr);
        if (hipGetLastError() == hipSuccess && hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(r, dr, bytes, hipMemcpyDeviceToHost) == hipSuccess)
            rc = 0;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dr) (void)hipFree(dr);
    return rc;
}

extern "C" int arith_check(int nwords, int op, int kclr, int n, const uint32_t* a,
                           const uint32_t* b, uint32_t* r)
{
    if (n < 1 || op < 0 || op > 4 || !a || !b || !r)
        return -1;
    switch (nwords) {
    case 1: return run<1>(op, kclr, n, a, b, r);
    case 2: return run<2>(op, kclr, n, a, b, r);
    case 3: return run<3>(op, kclr, n, a, b, r);
    case 4: return run<4>(op, kclr, n, a, b, r);
    case 6: return run<6>(op, kclr, n, a, b, r);
    case 8: return run<8>(op, kclr, n, a, b, r);
    case 12: return run<12>(op, kclr, n, a, b, r);
    case 16: return run<16>(op, kclr, n, a, b, r);
    case 24: return run<24>(op, kclr, n, a, b, r);
    case 32: return run<32>(op, kclr, n, a, b, r);
    default: return -1;
    }
}
