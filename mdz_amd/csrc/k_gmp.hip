/*
 * k_gmp.hip -- kernels of the GMP path: the loops of src/frac_mandel.c:55-80,
 * src/frac_burning_ship.c:58-83, src/frac_generalized_celtic.c:58-83 and
 * src/frac_variant.c:58-87 in GMP's truncating mpf arithmetic (mpf_gmp.h),
 * z0 and c from host-computed per-column/per-row values (the mpf chain of
 * src/fractal.c:298-343).
 *
 * Values are gf<L> of up to L limbs of storage; a tile's PREC + 1 (the limbs
 * its values may hold, dtile.gprec) is a run-time parameter, so one build per
 * storage bucket (MDZG_GMP_LIMBS) serves every precision up to it.  The three
 * launch forms mirror the RN kernels: one thread per pixel of a tile, the
 * persistent per-class queue, and resumable slices for the server.
 */
#include "k_escape.h"
#include "mpf_gmp.h"

using namespace mdzg::gmpf;

template <int L>
__device__ __forceinline__ void load_gf(gf<L>& v, const uint32_t* w)
{
#pragma unroll
    for (int i = 0; i < L; ++i) v.t[i] = (uint64_t)w[2 * i] | ((uint64_t)w[2 * i + 1] << 32);
    v.size = (int32_t)w[2 * L];
    v.exp = (int32_t)w[2 * L + 1];
    v.neg = w[2 * L + 2];
}

template <int L>
__device__ __forceinline__ void store_gf(uint32_t* w, const gf<L>& v)
{
#pragma unroll
    for (int i = 0; i < L; ++i) {
        w[2 * i] = (uint32_t)v.t[i];
        w[2 * i + 1] = (uint32_t)(v.t[i] >> 32);
    }
    w[2 * L] = (uint32_t)v.size;
    w[2 * L + 1] = (uint32_t)v.exp;
    w[2 * L + 2] = v.neg;
}

template <int L>
struct gmp_state {
    gf<L> wre, wim, wre2, wim2, cre, cim;
};

/* src/fractal.c:328-343: the squares of the start values (mpf_mul at p),
   c = z0 (Mandelbrot family) or the Julia constant */
template <int L>
__device__ __forceinline__ void gmp_init(gmp_state<L>& s, const mdzg_dtile& t, int row, int col)
{
    const int vs = t.vs, lr = t.gprec;
    load_gf<L>(s.wre, t.xs + (size_t)col * vs);
    load_gf<L>(s.wim, t.ys + (size_t)row * vs);
    gf_sqr<L>(s.wre2, s.wre, lr);
    gf_sqr<L>(s.wim2, s.wim, lr);
    if (t.family == 1) {
        load_gf<L>(s.cre, t.jc);
        load_gf<L>(s.cim, t.jc + vs);
    } else {
        s.cre = s.wre;
        s.cim = s.wim;
    }
}

/* one iteration of src/frac_mandel.c:66-79 (and siblings); true on escape */
template <int L>
__device__ __forceinline__ bool gmp_step(gmp_state<L>& s, int fractal, int32_t wz, int lr)
{
    gf<L> t1, t2;
    gf_mul<L>(t1, s.wre, s.wim, lr);                       /* t1 = wre * wim     */
    if (fractal == F_BURNING_SHIP)
        gf_abs<L>(t1);
    gf_mul2<L>(t2, t1, lr);                                /* t1 *= 2 (mul_ui)   */
    gf_addsub<L>(s.wim, t2, s.cim, 0u, lr);                /* wim = t1 + c_im    */
    gf_addsub<L>(t1, s.wre2, s.wim2, 1u, lr);              /* t1 = wre2 - wim2   */
    if (fractal == F_GENERALIZED_CELTIC || (fractal == F_VARIANT && (wz & 1)))
        gf_abs<L>(t1);
    gf_addsub<L>(s.wre, t1, s.cre, 0u, lr);                /* wre = t1 + c_re    */
    gf_sqr<L>(s.wim2, s.wim, lr);                          /* wim2 = wim^2       */
    gf_sqr<L>(s.wre2, s.wre, lr);                          /* wre2 = wre^2       */
    /* t1 = wim2 + wre2, mpf_cmp (t1, 4): the sum is formed only where the
       integer limbs leave the comparison open */
    const int q = gf_sum_gt4_quick<L>(s.wim2, s.wre2);
    bool esc = q > 0;
    if (gf_any(q < 0)) {
        gf_add<L>(t1, s.wim2, s.wre2, 0u, lr);
        esc = (q < 0) ? gf_gt4<L>(t1) : esc;
    }
    return esc;
}

template <int L>
__global__ __launch_bounds__(256) void k_gmp_tile(const mdzg_dtile* tp)
{
    const mdzg_dtile& t = *tp;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = (uint32_t)t.ncols * (uint32_t)t.nrows;
    if (idx >= npix)
        return;
    const int row = (int)(idx / (uint32_t)t.ncols), col = (int)(idx % (uint32_t)t.ncols);
    gmp_state<L> s;
    gmp_init<L>(s, t, row, col);
    const int fractal = t.fractal, lr = t.gprec;
    const int32_t depth = t.depth;
    uint32_t count = 0;
    for (int32_t wz = 1;; ++wz) {   /* ends at depth == INT32_MAX too */
        if (gmp_step<L>(s, fractal, wz, lr)) {
            count = (uint32_t)wz;
            break;
        }
        if (wz >= depth)
            break;
    }
    t.out[idx] = count;
}

template <int L>
__global__ __launch_bounds__(256) void k_gmp_persist(const mdzg_dtile* tiles, int ntiles,
                                                     uint64_t total, unsigned long long* counter,
                                                     const volatile int* cancel, int chunk)
{
    bool active = false;
    int fractal = 0, lr = 0;
    uint32_t round = 0;
    int32_t wz = 0, depth = 0;
    uint32_t* out = nullptr;
    gmp_state<L> s;
    bool drained = false;
    wave_run run = { 0, 0u };
    /* no more pixels than lanes: each wave starts with its own 64 slots as
       its run (taken without an atomic) and never claims again */
    const bool by_lane = total <= (uint64_t)gridDim.x * blockDim.x;
    if (by_lane) {
        run.base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
        run.left = 64u;
    }
    for (;;) {
        if (!drained) {
            const bool need = !active;
            const uint64_t g = claim<0u>(need, counter, run);
            if (need && g < total) {
                const int ti = find_tile(tiles, ntiles, g);
                const mdzg_dtile& t = tiles[ti];
                const uint32_t local = (uint32_t)(g - t.pix_begin);
                const int row = (int)(local / (uint32_t)t.ncols);
                const int col = (int)(local % (uint32_t)t.ncols);
                gmp_init<L>(s, t, row, col);
                fractal = t.fractal;
                lr = t.gprec;
                depth = t.depth;
                out = t.out + local;
                wz = 0;
                active = true;
                /* wait for the tile's fields here, not where the paths join
                   (there the wait would include the count stores in flight) */
                asm volatile("" ::"v"(fractal), "v"(lr), "v"(depth), "v"(out));
            }
            drained = by_lane || __any(need && g >= total);
        }
        if (!__any(active))
            break;
        const int stop = ((++round & 15) == 0) ? read_cancel(cancel) : 0;
        for (int it = 0; it < chunk; ++it) {
            if (active) {
                ++wz;
                const bool esc = gmp_step<L>(s, fractal, wz, lr);
                if (esc || wz >= depth) {
                    put_count(out, esc ? (uint32_t)wz : 0u);
                    active = false;
                }
            }
        }
        if (stop)
            break;
    }
}

/* Resumable slice (serving): see k_rn_slice. */
template <int L>
__global__ __launch_bounds__(256) void k_gmp_slice(const mdzg_dtile* tiles, int ntiles,
                                                   uint64_t total, unsigned long long* counter,
                                                   int budget,
    int lane_budget)
{
    bool active = false;
    int fractal = 0, lr = 0;
    int32_t wz = 0, depth = 0;
    int32_t left = 0;           /* this pixel's iterations left in the slice */
    int32_t lb = lane_budget;   /* this lane's iterations left in the slice */
    uint32_t local = 0;
    const mdzg_dtile* tp = tiles;
    gmp_state<L> s;
    bool drained = false;
    wave_run run = { 0, 0u };
    for (;;) {
        if (!drained) {
            const bool need = !active && lb > 0;
            const uint64_t g = claim<0u>(need, counter, run);
            if (need && g < total) {
                const int ti = find_tile(tiles, ntiles, g);
                tp = &tiles[ti];
                local = (uint32_t)(g - tp->pix_begin);
                if (tp->out[local] == 0xffffffffu) {
                    const int row = (int)(local / (uint32_t)tp->ncols);
                    const int col = (int)(local % (uint32_t)tp->ncols);
                    const int vs = tp->vs;
                    wz = tp->wzs[local];
                    if (wz == 0) {
                        gmp_init<L>(s, *tp, row, col);
                    } else {
                        const uint32_t* st = tp->state + (size_t)local * 4 * vs;
                        load_gf<L>(s.wre, st);
                        load_gf<L>(s.wim, st + vs);
                        load_gf<L>(s.wre2, st + 2 * vs);
                        load_gf<L>(s.wim2, st + 3 * vs);
                        if (tp->family == 1) {
                            load_gf<L>(s.cre, tp->jc);
                            load_gf<L>(s.cim, tp->jc + vs);
                        } else {
                            load_gf<L>(s.cre, tp->xs + (size_t)col * vs);
                            load_gf<L>(s.cim, tp->ys + (size_t)row * vs);
                        }
                    }
                    fractal = tp->fractal;
                    lr = tp->gprec;
                    depth = tp->depth;
                    left = budget;
                    active = true;
                }
            }
            drained = __any(need && g >= total) || !__any(lb > 0);
        }
        if (!__any(active)) {
            if (drained)
                break;
            continue;
        }
        if (active) {
            ++wz;
            --left;
            --lb;
            const bool esc = gmp_step<L>(s, fractal, wz, lr);
            if (esc || wz >= depth) {
                put_count(tp->out + local, esc ? (uint32_t)wz : 0u);
                atomicSub(tp->remaining, 1u);
                active = false;
            } else if (left == 0 || lb == 0) {
                const int vs = tp->vs;
                uint32_t* st = tp->state + (size_t)local * 4 * vs;
                store_gf<L>(st, s.wre);
                store_gf<L>(st + vs, s.wim);
                store_gf<L>(st + 2 * vs, s.wre2);
                store_gf<L>(st + 3 * vs, s.wim2);
                tp->wzs[local] = wz;
                active = false;
            }
        }
    }
}

extern "C" void mdzg_kget_gmp(int bucket, const void** tile, const void** persist,
                              const void** slice)
{
    switch (bucket) {
    case 0:
        *tile = (const void*)k_gmp_tile<4>;
        *persist = (const void*)k_gmp_persist<4>;
        *slice = (const void*)k_gmp_slice<4>;
        break;
    case 1:
        *tile = (const void*)k_gmp_tile<6>;
        *persist = (const void*)k_gmp_persist<6>;
        *slice = (const void*)k_gmp_slice<6>;
        break;
    case 2:
        *tile = (const void*)k_gmp_tile<10>;
        *persist = (const void*)k_gmp_persist<10>;
        *slice = (const void*)k_gmp_slice<10>;
        break;
    default:
        *tile = (const void*)k_gmp_tile<18>;
        *persist = (const void*)k_gmp_persist<18>;
        *slice = (const void*)k_gmp_slice<18>;
        break;
    }
}
