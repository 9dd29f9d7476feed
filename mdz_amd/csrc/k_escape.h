#ifndef MDZG_K_ESCAPE_H
#define MDZG_K_ESCAPE_H
/*
 * k_escape.h -- gfx950 escape-time kernels (included by k_fp64.hip and, once
 * per word count N, by k_rn.hip).
 *
 * Two arithmetics:
 *   FP64 : the loop of src/frac_mandel.c:5-21 (and siblings) in IEEE double,
 *          c from the double restatement of src/fractal.c:61-72.  Built with
 *          -ffp-contract=off (and the pragma below) so every operation is one
 *          IEEE rounding in C expression order, as the CPU double loop does.
 *   RN<N>: the MPFR loop of src/frac_mandel.c:33-51 (and siblings) with every
 *          operation rounded to nearest-even at the tile's precision p <= 32N
 *          (mp_rn.h), z0 and c from host-computed per-column/per-row values.
 *
 * Two launch forms per arithmetic:
 *   *_tile    : one thread per pixel of one tile (the one-launch-per-tile form);
 *   *_persist : a grid sized to the resident capacity of the GPU; each wave
 *               keeps 64 pixels in flight and refills lanes whose pixel has
 *               finished from a device queue (one 64-bit counter over the flat
 *               pixel space of all tiles of the class), so tiles of uneven
 *               size, depth and formula share one launch and lanes do not idle
 *               behind the slowest pixel of their wave.  Every wave leaves once
 *               the queue is drained and its lanes are done, or when the host
 *               raises the cancel flag.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mdzg_dev.h"
#include "mp_rn.h"

#pragma clang fp contract(off)

using namespace mdzg;

enum { F_MANDELBROT = 0, F_BURNING_SHIP = 1, F_GENERALIZED_CELTIC = 2, F_VARIANT = 3 };

/* ------------------------------------------------------------------ */
/* shared helpers                                                       */
/* ------------------------------------------------------------------ */
__device__ __forceinline__ int find_tile(const mdzg_dtile* tiles, int ntiles, uint64_t g)
{
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid].pix_begin <= g)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

/* Counts are stored and the cancel flag is read through global (address
   space 1) pointers.  Through a generic pointer they compile to flat
   instructions, which count against the LDS counter as well as the memory
   counter, so each lane shuffle of the next refill round would wait until the
   previous round's count stores had reached memory. */
typedef __attribute__((address_space(1))) uint32_t g_u32;
typedef __attribute__((address_space(1))) const volatile int g_cflag;

__device__ __forceinline__ void put_count(uint32_t* p, uint32_t v)
{
    *(g_u32*)p = v;
}

__device__ __forceinline__ int read_cancel(const volatile int* p)
{
    return *(g_cflag*)p;
}

/* Queue slots are claimed per wave.  A wave whose lanes need n slots takes
   what is left of its run first; the rest comes from one atomic on the
   shared counter, which takes max(missing, RUN) consecutive slots, the
   surplus kept as the wave's run for later refills.  For the cheap loops
   (FP64, RN of 1-2 words) a frame of quickly escaping pixels refills lanes
   constantly, and one atomic per refill from every wave made the shared
   counter the bottleneck: RUN = 64.  For the wide RN and GMP loops a pixel
   costs far more than an atomic, and a reserved run would lengthen the
   frame's tail by a pixel's time: RUN = 0, one slot per lane as it frees. */
struct wave_run {
    uint64_t base;      /* next slot of the run (the same on every lane) */
    uint32_t left;      /* slots left in the run */
};

__device__ __forceinline__ uint64_t claim_len(bool need, unsigned long long* counter, wave_run& w,
                                              uint32_t RUN)
{
    const uint64_t m = __ballot(need);
    uint64_t g = ~0ull;
    if (m) {
        const int lane = __lane_id();
        const uint32_t n = (uint32_t)__popcll(m);
        const uint32_t below = (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
        if (w.left < n) {
            const uint32_t miss = n - w.left;
            const uint32_t take = miss > RUN ? miss : RUN;
            const int leader = __ffsll((long long)m) - 1;
            unsigned long long nb = 0;
            if (lane == leader)
                nb = atomicAdd(counter, (unsigned long long)take);
            nb = __shfl(nb, leader);
            g = below < w.left ? w.base + below : nb + (below - w.left);
            w.base = nb + miss;
            w.left = take - miss;
        } else {
            g = w.base + below;
            w.base += n;
            w.left -= n;
        }
        if (!need)
            g = ~0ull;
    }
    return g;
}

template <uint32_t RUN>
__device__ __forceinline__ uint64_t claim(bool need, unsigned long long* counter, wave_run& w)
{
    return claim_len(need, counter, w, RUN);
}

/* run length for a launch of `total` slots: a K-th of each wave's even share
   of the queue, at least 64 and at most CAP.  Long runs save atomics on the
   shared counter (the bottleneck of frames whose pixels escape within a few
   iterations); capping a run at a fraction of the share bounds what one wave
   can still hold when the queue drains, and small launches keep runs of 64 so
   that every wave gets work. */
#ifndef MDZG_RUN_K
#define MDZG_RUN_K 4u
#endif
#ifndef MDZG_RUN_CAP
#define MDZG_RUN_CAP 256u
#endif
__device__ __forceinline__ uint32_t share_run(uint64_t total)
{
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x >> 6);
    const uint64_t r = total / (waves * MDZG_RUN_K);
    return r < 64u ? 64u : (r > MDZG_RUN_CAP ? MDZG_RUN_CAP : (uint32_t)r);
}

/* run length by loop: FP64 and RN up to 2 words claim in runs of 64 */
template <int N>
constexpr uint32_t claim_run()
{
    return N <= 2 ? 64u : 0u;
}

#ifdef MDZG_FP64_KERNELS
/* ------------------------------------------------------------------ */
/* FP64                                                                 */
/* ------------------------------------------------------------------ */
struct fp64_state {
    double wre, wim, wre2, wim2, cre, cim;
};

__device__ __forceinline__ void fp64_init(fp64_state& s, const mdzg_dtile& t, int row, int col)
{
    const int ix = t.col0 + col;
    const int line = t.line0 + row;
    const double W = (double)t.W;
    const double y = t.ymax - (t.width / W) * (double)line;
    const double x = (ix / W) * t.width + t.xmin;
    s.wre = x;
    s.wim = y;
    s.wre2 = x * x;
    s.wim2 = y * y;
    if (t.family == 1) {
        s.cre = t.jre;
        s.cim = t.jim;
    } else {
        s.cre = x;
        s.cim = y;
    }
}

/* one iteration; returns true when |z|^2 > 4 (src/frac_mandel.c:13-18) */
__device__ __forceinline__ bool fp64_step(fp64_state& s, int fractal, int32_t wz)
{
    if (fractal == F_BURNING_SHIP)
        s.wim = 2.0 * fabs(s.wre * s.wim) + s.cim;
    else
        s.wim = 2.0 * s.wre * s.wim + s.cim;
    if (fractal == F_GENERALIZED_CELTIC || (fractal == F_VARIANT && (wz & 1)))
        s.wre = fabs(s.wre2 - s.wim2) + s.cre;
    else
        s.wre = s.wre2 - s.wim2 + s.cre;
    s.wim2 = s.wim * s.wim;
    s.wre2 = s.wre * s.wre;
    return s.wim2 + s.wre2 > 4.0;
}

__global__ __launch_bounds__(256) void k_fp64_tile(const mdzg_dtile* tp)
{
    const mdzg_dtile t = *tp;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = (uint32_t)t.ncols * (uint32_t)t.nrows;
    if (idx >= npix)
        return;
    const int row = (int)(idx / (uint32_t)t.ncols), col = (int)(idx % (uint32_t)t.ncols);
    fp64_state s;
    fp64_init(s, t, row, col);
    uint32_t count = 0;
    /* the exit test follows the step, so depth == INT32_MAX ends too */
    for (int32_t wz = 1;; ++wz) {
        if (fp64_step(s, t.fractal, wz)) {
            count = (uint32_t)wz;
            break;
        }
        if (wz >= t.depth)
            break;
    }
    t.out[idx] = count;
}

/* v + 1 on the lanes set in the lane mask m, v elsewhere: one
   v_addc_co_u32 with m as the carry-in */
__device__ __forceinline__ int32_t add_lanes(int32_t v, uint64_t m)
{
    int32_t r;
    uint64_t co;
    asm("v_addc_co_u32 %0, %1, 0, %2, %3" : "=v"(r), "=s"(co) : "v"(v), "s"(m));
    return r;
}

/* One iteration of formula F (a compile-time constant) for every lane of the
   wave, without branches.  live is the lane mask of pixels still iterating:
   their wz counts the iteration, and a lane that escapes leaves the mask, so
   its wz stops at its count.  Other lanes compute on their last values, which
   no one reads.  Per iteration that is the formula's FP64 operations, one
   compare into a scalar mask and one add; the per-lane branches of the
   generic loop cost more scalar exec-mask instructions per iteration than the
   loop has VALU instructions.  With DEPTH a lane also leaves the mask at its
   depth (one more compare), and escm collects the lanes that escaped. */
template <int F, bool DEPTH>
__device__ __forceinline__ void fp64_iter(fp64_state& s, int32_t& wz, uint64_t& live,
                                          uint64_t& escm, int32_t depth)
{
    const bool esc = fp64_step(s, F, wz + 1);
    const uint64_t e = __builtin_amdgcn_ballot_w64(esc);
    wz = add_lanes(wz, live);
    if (DEPTH) {
        escm |= e & live;
        live &= ~(e | __builtin_amdgcn_ballot_w64(wz >= depth));
    } else {
        live &= ~e;
    }
}

/* Up to n iterations of formula F, fewer once no lane is left iterating;
   escm gets the lanes that escaped.  Without DEPTH the caller has checked
   that no live lane reaches its depth within n iterations. */
template <int F, bool DEPTH>
__device__ __forceinline__ void fp64_run(fp64_state& s, int32_t& wz, uint64_t& live,
                                         uint64_t& escm, int32_t depth, int n)
{
    const uint64_t start = live;
    int it = 0;
    for (; it + 4 <= n && live != 0; it += 4) {
        fp64_iter<F, DEPTH>(s, wz, live, escm, depth);
        fp64_iter<F, DEPTH>(s, wz, live, escm, depth);
        fp64_iter<F, DEPTH>(s, wz, live, escm, depth);
        fp64_iter<F, DEPTH>(s, wz, live, escm, depth);
    }
    for (; it < n && live != 0; ++it)
        fp64_iter<F, DEPTH>(s, wz, live, escm, depth);
    if (!DEPTH)
        escm = start & ~live;
}

/* The latency form of an iteration, for launches whose pixels all fit the
   device's lanes: each wave then runs alone on its SIMD, and the time of an
   iteration is the length of its dependent chain.  The lane's live flag (lv)
   and escape flag (ev) stay in vector registers, so that no scalar
   instruction stands between one iteration's compare and the next
   iteration's arithmetic (in the mask form, the count add of the next
   iteration waits for the scalar update of the mask, which waits for the
   compare); one more VALU instruction per iteration. */
template <int F, bool DEPTH>
__device__ __forceinline__ void fp64_iter_v(fp64_state& s, int32_t& wz, int32_t& lv, int32_t& ev,
                                            int32_t depth)
{
    const bool esc = fp64_step(s, F, wz + 1);
    wz += lv;
    if (DEPTH) {
        ev = esc ? (ev | lv) : ev;
        lv = (esc || wz >= depth) ? 0 : lv;
    } else {
        lv = esc ? 0 : lv;
    }
}

/* fp64_run in the latency form: the wave tests for an exit every 4
   iterations, as fp64_run does */
template <int F, bool DEPTH>
__device__ __forceinline__ void fp64_run_v(fp64_state& s, int32_t& wz, uint64_t& live,
                                           uint64_t& escm, int32_t depth, int n)
{
    const uint64_t start = live;
    int32_t lv = (int32_t)((live >> __lane_id()) & 1u), ev = 0;
    int it = 0;
    for (; it + 4 <= n && __builtin_amdgcn_ballot_w64(lv != 0) != 0; it += 4) {
        fp64_iter_v<F, DEPTH>(s, wz, lv, ev, depth);
        fp64_iter_v<F, DEPTH>(s, wz, lv, ev, depth);
        fp64_iter_v<F, DEPTH>(s, wz, lv, ev, depth);
        fp64_iter_v<F, DEPTH>(s, wz, lv, ev, depth);
    }
    for (; it < n && __builtin_amdgcn_ballot_w64(lv != 0) != 0; ++it)
        fp64_iter_v<F, DEPTH>(s, wz, lv, ev, depth);
    live = __builtin_amdgcn_ballot_w64(lv != 0);
    escm = DEPTH ? (escm | __builtin_amdgcn_ballot_w64(ev != 0)) : (start & ~live);
}

template <bool DEPTH, bool LAT>
__device__ __forceinline__ void fp64_run_formula(int f, fp64_state& s, int32_t& wz, uint64_t& live,
                                                 uint64_t& escm, int32_t depth, int n)
{
    if (LAT) {
        switch (f) {
        case F_MANDELBROT: fp64_run_v<F_MANDELBROT, DEPTH>(s, wz, live, escm, depth, n); break;
        case F_BURNING_SHIP: fp64_run_v<F_BURNING_SHIP, DEPTH>(s, wz, live, escm, depth, n); break;
        case F_GENERALIZED_CELTIC:
            fp64_run_v<F_GENERALIZED_CELTIC, DEPTH>(s, wz, live, escm, depth, n);
            break;
        default: fp64_run_v<F_VARIANT, DEPTH>(s, wz, live, escm, depth, n); break;
        }
        return;
    }
    switch (f) {
    case F_MANDELBROT: fp64_run<F_MANDELBROT, DEPTH>(s, wz, live, escm, depth, n); break;
    case F_BURNING_SHIP: fp64_run<F_BURNING_SHIP, DEPTH>(s, wz, live, escm, depth, n); break;
    case F_GENERALIZED_CELTIC:
        fp64_run<F_GENERALIZED_CELTIC, DEPTH>(s, wz, live, escm, depth, n);
        break;
    default: fp64_run<F_VARIANT, DEPTH>(s, wz, live, escm, depth, n); break;
    }
}

/* One run of `chunk` iterations for the wave.  No live lane passes its
   depth in the first rem iterations (rem >= 1: an active lane has
   wz < depth): those run without the depth test; lanes that reached their
   depth there leave the mask (count 0 unless they escaped, which already
   took them out), and the rest of the run has the test. */
template <bool LAT>
__device__ __forceinline__ void fp64_chunk(int f0, fp64_state& s, int32_t& wz, uint64_t& live,
                                           uint64_t& escm, int32_t depth, int chunk, int rem)
{
    if (rem >= chunk) {
        fp64_run_formula<false, LAT>(f0, s, wz, live, escm, depth, chunk);
    } else {
        fp64_run_formula<false, LAT>(f0, s, wz, live, escm, depth, rem);
        live &= ~__builtin_amdgcn_ballot_w64(wz >= depth);
        fp64_run_formula<true, LAT>(f0, s, wz, live, escm, depth, chunk - rem);
    }
}

__global__ __launch_bounds__(256) void k_fp64_persist(const mdzg_dtile* tiles, int ntiles,
                                                      uint64_t total,
                                                      unsigned long long* counter,
                                                      const volatile int* cancel, int chunk)
{
    bool active = false;
    int fractal = 0;
    uint32_t round = 0;
    int32_t wz = 0, depth = 0;
    uint32_t* out = nullptr;
    fp64_state s;
    s.wre = s.wim = s.wre2 = s.wim2 = s.cre = s.cim = 0.0;
    bool drained = false;
    wave_run run = { 0, 0u };
    const uint32_t run_len = share_run(total);
    /* a launch with no more pixels than lanes (the 240x180 frame) takes them
       by lane index: no queue atomic to start, none to find the queue empty */
    const bool by_lane = total <= (uint64_t)gridDim.x * blockDim.x;
    for (;;) {
        if (!drained) {
            const bool need = !active;
            const uint64_t g = by_lane ? (uint64_t)blockIdx.x * blockDim.x + threadIdx.x
                                       : claim_len(need, counter, run, run_len);
            if (need && g < total) {
                const int ti = find_tile(tiles, ntiles, g);
                const mdzg_dtile& t = tiles[ti];
                const uint32_t local = (uint32_t)(g - t.pix_begin);
                const int row = (int)(local / (uint32_t)t.ncols);
                const int col = (int)(local % (uint32_t)t.ncols);
                fp64_init(s, t, row, col);
                fractal = t.fractal;
                depth = t.depth;
                out = t.out + local;
                wz = 0;
                active = true;
                /* the tile's fields are waited for here, on the refill path:
                   a wait where the paths join would also wait, every round,
                   for the count stores still in flight */
                asm volatile("" ::"v"(fractal), "v"(depth), "v"(out));
            }
            /* the counter only grows: once one lane reads past the end the wave
               stops claiming, so no atomics while its last lanes finish */
            drained = by_lane || __any(need && g >= total);
        }
        if (!__any(active))
            break;
        /* cancel flag (device memory, written by the host on another stream):
           loaded before the chunk, tested after it, every 16th round */
        const int stop = ((++round & 15) == 0) ? read_cancel(cancel) : 0;
        /* one formula for the whole wave (the usual case): a branch-free run
           of `chunk` iterations.  When no pixel of the wave can reach its
           depth within it (the wave minimum of the remaining depths), the
           run skips the per-lane depth test.  Cutting the run at that
           minimum instead made the runs of waves whose pixels started at
           different times short: 1.60 against 1.82 T pixel-iter/s on a
           3840x2160 frame of the whole set (centre -0.75, width 3, depth
           1000). */
        const uint64_t am = __ballot(active);
        const int lead = __ffsll((long long)am) - 1;
        /* a scalar, so that the formula switch below is a uniform branch and
           the lane masks stay in scalar registers across it */
        const int f0 = __builtin_amdgcn_readfirstlane(__shfl(fractal, lead));
        if (__all(!active || fractal == f0)) {
            /* the wave minimum, read back as a scalar so that the compiler
               keeps the run loops' bounds, and the lane masks they update,
               uniform (scalar registers) */
            const int rem = __builtin_amdgcn_readfirstlane(
                __ockl_wfred_min_i32(active ? depth - wz : 0x7fffffff));
            /* add_lanes' asm reads live: a scalar copy of the ballot (later
               iterations update it with scalar instructions) */
            uint64_t live = scalar_copy(am), escm = 0;
            if (by_lane)
                fp64_chunk<true>(f0, s, wz, live, escm, depth, chunk, rem);
            else
                fp64_chunk<false>(f0, s, wz, live, escm, depth, chunk, rem);
            const bool esc = (escm >> __lane_id()) & 1u;
            if (active && (esc || wz >= depth)) {
                put_count(out, esc ? (uint32_t)wz : 0u);
                active = false;
            }
        } else {
            for (int it = 0; it < chunk; ++it) {
                if (active) {
                    ++wz;
                    const bool esc = fp64_step(s, fractal, wz);
                    if (esc || wz >= depth) {
                        put_count(out, esc ? (uint32_t)wz : 0u);
                        active = false;
                    }
                }
            }
        }
        if (stop)
            break;
    }
}

/* Resumable slice (serving).  Every lane has `budget` iterations for the
   slice.  It claims the tiles' unfinished pixels in queue order (the server
   puts the requests with the least work left first) and runs each until it
   finishes -- its count written, its request's remaining counter
   decremented -- or until the lane's budget is spent, when it saves z, z^2
   and wz.  Pixels the slice does not reach keep their state for the next.
   A slice therefore lasts about `budget` iterations however many pixels are
   active, and a small request at the front of the queue advances as fast as
   it would alone. */
__global__ __launch_bounds__(256) void k_fp64_slice(const mdzg_dtile* tiles, int ntiles,
                                                    uint64_t total,
                                                    unsigned long long* counter, int budget,
    int lane_budget)
{
    bool active = false;
    int fractal = 0;
    int32_t wz = 0, depth = 0;
    int32_t left = 0;           /* this pixel's iterations left in the slice */
    int32_t lb = lane_budget;   /* this lane's iterations left in the slice */
    uint32_t local = 0;
    const mdzg_dtile* tp = tiles;
    fp64_state s;
    s.wre = s.wim = s.wre2 = s.wim2 = s.cre = s.cim = 0.0;
    bool drained = false;
    wave_run run = { 0, 0u };
    for (;;) {
        if (!drained) {
          const bool need = !active && lb > 0;
          const uint64_t g = claim<64u>(need, counter, run);
          if (need && g < total) {
            const int ti = find_tile(tiles, ntiles, g);
            tp = &tiles[ti];
            local = (uint32_t)(g - tp->pix_begin);
            if (tp->out[local] == 0xffffffffu) {
                const int row = (int)(local / (uint32_t)tp->ncols);
                const int col = (int)(local % (uint32_t)tp->ncols);
                fp64_init(s, *tp, row, col);
                wz = tp->wzs[local];
                if (wz > 0) {
                    const double* st = (const double*)tp->state + 4 * (size_t)local;
                    s.wre = st[0];
                    s.wim = st[1];
                    s.wre2 = st[2];
                    s.wim2 = st[3];
                }
                fractal = tp->fractal;
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
        for (int it = 0; it < 32; ++it) {
            if (active) {
                ++wz;
                --left;
                --lb;
                const bool esc = fp64_step(s, fractal, wz);
                if (esc || wz >= depth) {
                    put_count(tp->out + local, esc ? (uint32_t)wz : 0u);
                    atomicSub(tp->remaining, 1u);
                    active = false;
                } else if (left == 0 || lb == 0) {
                    double* st = (double*)tp->state + 4 * (size_t)local;
                    st[0] = s.wre;
                    st[1] = s.wim;
                    st[2] = s.wre2;
                    st[3] = s.wim2;
                    tp->wzs[local] = wz;
                    active = false;
                }
            }
        }
    }
}

#endif /* MDZG_FP64_KERNELS */

/* ------------------------------------------------------------------ */
/* RN<N>                                                                */
/* ------------------------------------------------------------------ */
template <int N>
__device__ __forceinline__ void load_mp(mp<N>& v, const uint32_t* p)
{
#pragma unroll
    for (int j = 0; j < N; ++j) v.m[j] = p[j];
    v.e = (int32_t)p[N];
    v.s = p[N + 1];
}

/* minimum waves per SIMD the RN kernels are compiled for: small N keep
   several waves resident, large N get the register file to themselves */
template <int N, bool FULL>
constexpr int rn_min_waves()
{
    return (N >= 32 && !FULL) ? 1 : (N >= 12 ? 2 : 4);
}

/* the slice kernel keeps its resume and save paths live inside the
   iteration loop, so it gets a larger share of the register file than the
   other kernels from 8 words up: two waves per SIMD (one for the ragged
   32-word class, as the other kernels) */
template <int N, bool FULL>
constexpr int rn_min_waves_slice()
{
    return (N >= 32 && !FULL) ? 1 : (N >= 8 ? 2 : rn_min_waves<N, FULL>());
}

/* c per lane: registers for small N; for N >= 24 the block's LDS (word w of
   lane t at [w][t], conflict-free), so the register file holds z and z^2 only
   and two waves fit per SIMD */
template <int N>
constexpr bool c_in_lds()
{
    return N >= 24;
}

template <int N, bool LDS>
struct c_src;

template <int N>
struct c_src<N, false> {
    mp<N> re, im;
    __device__ __forceinline__ void init(uint32_t*) {}
    __device__ __forceinline__ void put(const mp<N>& r, const mp<N>& i)
    {
        re = r;
        im = i;
    }
    __device__ __forceinline__ const mp<N>& get_re() const { return re; }
    __device__ __forceinline__ const mp<N>& get_im() const { return im; }
};

template <int N>
struct c_src<N, true> {
    uint32_t* base;   /* this lane's column of the block's LDS area */
    __device__ __forceinline__ void init(uint32_t* lds) { base = lds + threadIdx.x; }
    __device__ __forceinline__ void put(const mp<N>& r, const mp<N>& i)
    {
#pragma unroll
        for (int j = 0; j < N; ++j) {
            base[j * 256] = r.m[j];
            base[(N + 2 + j) * 256] = i.m[j];
        }
        base[N * 256] = (uint32_t)r.e;
        base[(N + 1) * 256] = r.s;
        base[(2 * N + 2) * 256] = (uint32_t)i.e;
        base[(2 * N + 3) * 256] = i.s;
    }
    __device__ __forceinline__ mp<N> get(int off) const
    {
        mp<N> v;
#pragma unroll
        for (int j = 0; j < N; ++j) v.m[j] = base[(off + j) * 256];
        v.e = (int32_t)base[(off + N) * 256];
        v.s = base[(off + N + 1) * 256];
        return v;
    }
    __device__ __forceinline__ mp<N> get_re() const { return get(0); }
    __device__ __forceinline__ mp<N> get_im() const { return get(N + 2); }
};

template <int N>
struct rn_state {
    mp<N> wre, wim, wre2, wim2;
    c_src<N, c_in_lds<N>()> c;
    bool lazy;   /* wim2 holds only a bound on its exponent (rn_step, LAZY) */
};

template <int N, bool FULL>
__device__ __forceinline__ void rn_init(rn_state<N>& s, const mdzg_dtile& t, int row, int col)
{
    const int k = FULL ? 0 : t.kclr;
    load_mp<N>(s.wre, t.xs + (size_t)col * MDZG_VS(N));
    load_mp<N>(s.wim, t.ys + (size_t)row * MDZG_VS(N));
    mul_rn<N, true>(s.wre2, s.wre, s.wre, k);      /* src/fractal.c:188 */
    mul_rn<N, true>(s.wim2, s.wim, s.wim, k);      /* src/fractal.c:171 */
    s.lazy = false;
    if (t.family == 1) {
        mp<N> jr, ji;
        load_mp<N>(jr, t.jc);
        load_mp<N>(ji, t.jc + MDZG_VS(N));
        s.c.put(jr, ji);
    } else {
        s.c.put(s.wre, s.wim);
    }
}

/* one iteration of src/frac_mandel.c:36-48 (and siblings); true on escape.

   LAZY (word-aligned kernels from 12 words; the squares' difference returns
   its minuend when every lane's subtrahend lies 32N + 2 bits or more below
   it): wim2 enters the next iteration only through that
   difference and this iteration's escape test.  When wim^2 is that far below
   wre2 on every lane of the wave, as near the real axis in every iteration,
   neither reads its bits: RN(wre2 - wim2) = wre2 and RN(wim2 + wre2) = wre2
   for any value below a quarter of wre2's last place, and the escape test's
   double estimate does not see it either.  The wave then keeps only an upper
   bound on wim2's exponent, 2 e(wim) + 1 (rounding can carry wim^2 up to the
   next power of two), and marks the lanes lazy.  At the start of the next
   iteration, while wim still holds the value those squares square, a wave
   whose lanes do not all pass the difference's shortcut (a lane refilled
   with a new pixel, or a growing wim) computes the exact squares of its lazy
   lanes first.  Counts and every value the loop reads are those of the
   exact loop.  The decision needs only exponents: e(wim^2) <= 2 e(wim) + 1
   and e(wre^2) >= 2 e(wre) - 1.  Near the real axis it took 1024-bit frames
   from 12.4 to 15.6 G pixel-iter/s; off the axis its votes and state cost
   2-3 % from 12 words and 3-9 % at 6 and 8, which keep the plain loop. */
template <int N, bool FULL>
__device__ constexpr bool rn_lazy() { return FULL && N >= 12; }

template <int N, bool FULL, bool LAZY = false>
__device__ __forceinline__ bool rn_step(rn_state<N>& s, int fractal, int32_t wz, int kclr)
{
    const int k = FULL ? 0 : kclr;
    constexpr bool LZ = LAZY && rn_lazy<N, FULL>();
    constexpr int32_t GAP = 32 * N + 2;
    if constexpr (LZ) {
        if (__builtin_expect(wave_any(s.lazy) &&
                                 !wave_all((int64_t)s.wre2.e - s.wim2.e >= GAP), 0)) {
            mp<N> sq;
            mul_rn<N, true>(sq, s.wim, s.wim, k);
            if (s.lazy)
                s.wim2 = sq;
            s.lazy = false;
        }
    }
    mp<N> t;
    mul_rn<N, false>(t, s.wre, s.wim, k);                    /* t1 = wre*wim     */
    t.s = (fractal == F_BURNING_SHIP) ? 0u : t.s;            /* BS: |t1|         */
    twice<N>(t);                                             /* t1 *= 2          */
    add_rn<N>(s.wim, t, s.c.get_im(), 0u, k);                /* wim = t1 + c_im  */
    add_rn<N, SUB_NONNEG, FULL>(t, s.wre2, s.wim2, 1u, k);   /* t1 = wre2 - wim2 */
    const bool ab = fractal == F_GENERALIZED_CELTIC || (fractal == F_VARIANT && (wz & 1));
    t.s = ab ? 0u : t.s;                                     /* Celtic: |t1|     */
    add_rn<N>(s.wre, t, s.c.get_re(), 0u, k);                /* wre = t1 + c_re  */
    if constexpr (LZ) {
        /* e(wim^2) <= 2 e(wim) + 1 and e(wre^2) >= 2 e(wre) - 1 */
        const int64_t eb = 2 * (int64_t)s.wim.e + 1;
        const bool far = !is_zero<N>(s.wim) & !is_zero<N>(s.wre) &
                         (2 * (int64_t)s.wre.e - 1 - eb >= GAP);
        if (wave_all(far)) {
#pragma unroll
            for (int j = 0; j < N - 1; ++j) s.wim2.m[j] = 0;
            s.wim2.m[N - 1] = 0x80000000u;
            s.wim2.e = (int32_t)eb;
            s.wim2.s = 0;
            s.lazy = true;
        } else {
            mul_rn<N, true>(s.wim2, s.wim, s.wim, k);        /* wim2 = wim^2     */
            s.lazy = false;
        }
    } else {
        mul_rn<N, true>(s.wim2, s.wim, s.wim, k);            /* wim2 = wim^2     */
    }
    mul_rn<N, true>(s.wre2, s.wre, s.wre, k);                /* wre2 = wre^2     */
    /* RN(wim2 + wre2) > 4  <=>  wim2 + wre2 > 4 + 2^(2-p) */
    const double thresh = 4.0 + ldexp(1.0, 2 - (32 * N - k));
    return escaped<N>(s.wim2, s.wre2, thresh, k);
}

template <int N, bool FULL>
__global__ __launch_bounds__(256, (rn_min_waves<N, FULL>())) void k_rn_tile(const mdzg_dtile* tp)
{
    const mdzg_dtile& t = *tp;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = (uint32_t)t.ncols * (uint32_t)t.nrows;
    if (idx >= npix)
        return;
    const int row = (int)(idx / (uint32_t)t.ncols), col = (int)(idx % (uint32_t)t.ncols);
    __shared__ uint32_t c_lds[c_in_lds<N>() ? 2 * (N + 2) * 256 : 1];
    rn_state<N> s;
    s.c.init(c_lds);
    rn_init<N, FULL>(s, t, row, col);
    const int fractal = t.fractal, kclr = t.kclr;
    const int32_t depth = t.depth;
    uint32_t count = 0;
    for (int32_t wz = 1;; ++wz) {
        if (rn_step<N, FULL, true>(s, fractal, wz, kclr)) {
            count = (uint32_t)wz;
            break;
        }
        if (wz >= depth)
            break;
    }
    t.out[idx] = count;
}

/* Per-operation probe (measurement only, p = 32N).  Each lane takes its
   pixel of the tile, runs `warm` iterations of the escape loop so that its
   values carry the orbit's exponents and bit patterns (lanes of a wave
   neighbours, as in a render), then repeats one operation `reps` times in a
   dependent chain whose operands keep those statistics:
     MDZG_OPP_LOOP   the loop alone (the chain's register dependency only)
     MDZG_OPP_MUL    x = RN(x * y)
     MDZG_OPP_SQR    x = RN(x^2) with the exponent set back to 0
     MDZG_OPP_ADD    x = RN(x + c_re), then x = RN(x - c_re): two a rep
     MDZG_OPP_SUBNN  a = |RN(a - b)|, a and b the squares (the SUB_NONNEG form)
     MDZG_OPP_ESC    the escape test on the squares
     MDZG_OPP_STEP   the whole iteration
   Counter passes over the launches give each operation's instructions.  A
   word of the result goes to the tile's output so that nothing is removed;
   it is not a count. */
template <int N>
__global__ __launch_bounds__(256, (rn_min_waves<N, true>())) void k_rn_op_probe(
    const mdzg_dtile* tp, int op, int warm, int reps)
{
    const mdzg_dtile& t = *tp;
    const uint32_t idx = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t npix = (uint32_t)t.ncols * (uint32_t)t.nrows;
    if (idx >= npix)
        return;
    const int row = (int)(idx / (uint32_t)t.ncols), col = (int)(idx % (uint32_t)t.ncols);
    __shared__ uint32_t c_lds[c_in_lds<N>() ? 2 * (N + 2) * 256 : 1];
    rn_state<N> s;
    s.c.init(c_lds);
    rn_init<N, true>(s, t, row, col);
    const int fractal = t.fractal;
    for (int32_t wz = 1; wz <= warm; ++wz)
        (void)rn_step<N, true>(s, fractal, wz, 0);
    const double thresh = 4.0 + ldexp(1.0, 2 - 32 * N);
    uint32_t h = 0;
    mp<N> x = s.wre, y = s.wim;
    /* one loop per operation, so that no loop carries the other
       operations' values */
    switch (op) {
    case MDZG_OPP_MUL:
        for (int i = 0; i < reps; ++i)
            mul_rn<N, false>(x, x, y, 0);
        break;
    case MDZG_OPP_SQR:
        for (int i = 0; i < reps; ++i) {
            mul_rn<N, true>(x, x, x, 0);
            x.e = is_zero(x) ? x.e : 0;
        }
        break;
    case MDZG_OPP_ADD:
        for (int i = 0; i < reps; ++i) {
            add_rn<N>(x, x, s.c.get_re(), 0u, 0);
            add_rn<N>(x, x, s.c.get_re(), 1u, 0);
        }
        break;
    case MDZG_OPP_SUBNN:
        for (int i = 0; i < reps; ++i) {
            add_rn<N, SUB_NONNEG>(x, s.wre2, s.wim2, 1u, 0);
            x.s = 0;
            s.wre2 = x;
        }
        break;
    case MDZG_OPP_ESC:
        for (int i = 0; i < reps; ++i) {
            h += escaped<N>(s.wim2, s.wre2, thresh, 0) ? 1u : 0u;
            asm volatile("" : "+v"(s.wim2.m[0]));
        }
        break;
    case MDZG_OPP_STEP:
        for (int i = 0; i < reps; ++i)
            h += rn_step<N, true>(s, fractal, i, 0) ? 1u : 0u;
        break;
    default:
        for (int i = 0; i < reps; ++i)
            asm volatile("" : "+v"(x.m[0]));
        break;
    }
    t.out[idx] = h ^ x.m[N - 1] ^ s.wre.m[N - 1] ^ s.wre2.m[N - 1];
}

template <int N, bool FULL>
__global__ __launch_bounds__(256, (rn_min_waves<N, FULL>())) void k_rn_persist(const mdzg_dtile* tiles, int ntiles,
                                                    uint64_t total,
                                                    unsigned long long* counter,
                                                    const volatile int* cancel, int chunk)
{
    bool active = false;
    int fractal = 0, kclr = 0;
    uint32_t round = 0;
    int32_t wz = 0, depth = 0;
    uint32_t* out = nullptr;
    __shared__ uint32_t c_lds[c_in_lds<N>() ? 2 * (N + 2) * 256 : 1];
    rn_state<N> s;
    s.c.init(c_lds);
    bool drained = false;
    wave_run run = { 0, 0u };
    /* no more pixels than lanes: each wave starts with its own 64 slots as
       its run (taken without an atomic) and never claims again (a 640x480
       frame at 128 bits: +2.5 %).  Not from 24 words: the start-up code changed the register
       allocation of those 256-VGPR kernels, 6 % slower on the 1024-bit band,
       whose launches never fit their lanes. */
    const bool by_lane = N < 24 && total <= (uint64_t)gridDim.x * blockDim.x;
    if (by_lane) {
        run.base = (uint64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63u);
        run.left = 64u;
    }
    for (;;) {
        if (!drained) {
            const bool need = !active;
            const uint64_t g = claim<claim_run<N>()>(need, counter, run);
            if (need && g < total) {
                const int ti = find_tile(tiles, ntiles, g);
                const mdzg_dtile& t = tiles[ti];
                const uint32_t local = (uint32_t)(g - t.pix_begin);
                const int row = (int)(local / (uint32_t)t.ncols);
                const int col = (int)(local % (uint32_t)t.ncols);
                rn_init<N, FULL>(s, t, row, col);
                fractal = t.fractal;
                kclr = t.kclr;
                depth = t.depth;
                out = t.out + local;
                wz = 0;
                active = true;
                asm volatile("" ::"v"(fractal), "v"(kclr), "v"(depth), "v"(out));
            }
            drained = by_lane || __any(need && g >= total);
        }
        if (!__any(active))
            break;
        const int stop = ((++round & 15) == 0) ? read_cancel(cancel) : 0;
        for (int it = 0; it < chunk; ++it) {
            if (active) {
                ++wz;
                const bool esc = rn_step<N, FULL, true>(s, fractal, wz, kclr);
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

template <int N>
__device__ __forceinline__ void store_mp(uint32_t* p, const mp<N>& v)
{
#pragma unroll
    for (int j = 0; j < N; ++j) p[j] = v.m[j];
    p[N] = (uint32_t)v.e;
    p[N + 1] = v.s;
}

/* Resumable slice (serving), RN arithmetic: see k_fp64_slice. */
template <int N, bool FULL>
__global__ __launch_bounds__(256, (rn_min_waves_slice<N, FULL>())) void k_rn_slice(
    const mdzg_dtile* tiles, int ntiles, uint64_t total, unsigned long long* counter, int budget,
    int lane_budget)
{
    bool active = false;
    int fractal = 0, kclr = 0;
    int32_t wz = 0, depth = 0;
    int32_t left = 0;           /* this pixel's iterations left in the slice */
    int32_t lb = lane_budget;   /* this lane's iterations left in the slice */
    uint32_t local = 0;
    const mdzg_dtile* tp = tiles;
    __shared__ uint32_t c_lds[c_in_lds<N>() ? 2 * (N + 2) * 256 : 1];
    rn_state<N> s;
    s.c.init(c_lds);
    bool drained = false;
    wave_run run = { 0, 0u };
    /* the claim loop runs until the wave holds a pixel or the queue is
       drained, and the kernel has a single exit: with a `continue` back to
       the claim instead, the 1024-bit kernel needed more than the 256 VGPRs
       of two waves per SIMD */
    for (;;) {
        do {
          if (drained)
              break;
          const bool need = !active && lb > 0;
          const uint64_t g = claim<claim_run<N>()>(need, counter, run);
          if (need && g < total) {
            const int ti = find_tile(tiles, ntiles, g);
            tp = &tiles[ti];
            local = (uint32_t)(g - tp->pix_begin);
            if (tp->out[local] == 0xffffffffu) {
                const int row = (int)(local / (uint32_t)tp->ncols);
                const int col = (int)(local % (uint32_t)tp->ncols);
                wz = tp->wzs[local];
                if (wz == 0) {
                    rn_init<N, FULL>(s, *tp, row, col);
                } else {
                    const uint32_t* st = tp->state + (size_t)local * 4 * MDZG_VS(N);
                    load_mp<N>(s.wre, st);
                    load_mp<N>(s.wim, st + MDZG_VS(N));
                    load_mp<N>(s.wre2, st + 2 * MDZG_VS(N));
                    load_mp<N>(s.wim2, st + 3 * MDZG_VS(N));
                    s.lazy = false;
                    mp<N> cr, ci;
                    if (tp->family == 1) {
                        load_mp<N>(cr, tp->jc);
                        load_mp<N>(ci, tp->jc + MDZG_VS(N));
                    } else {
                        load_mp<N>(cr, tp->xs + (size_t)col * MDZG_VS(N));
                        load_mp<N>(ci, tp->ys + (size_t)row * MDZG_VS(N));
                    }
                    s.c.put(cr, ci);
                }
                fractal = tp->fractal;
                kclr = tp->kclr;
                depth = tp->depth;
                left = budget;
                active = true;
            }
          }
          drained = __any(need && g >= total) || !__any(lb > 0);
        } while (!__any(active));
        if (!__any(active))
            break;
        if (active) {
            ++wz;
            --left;
            --lb;
            const bool esc = rn_step<N, FULL, true>(s, fractal, wz, kclr);
            if (esc || wz >= depth) {
                put_count(tp->out + local, esc ? (uint32_t)wz : 0u);
                atomicSub(tp->remaining, 1u);
                active = false;
            } else if (left == 0 || lb == 0) {
                /* a lazy wim2 is only a bound: the saved state holds the
                   exact square (wim is the value it squares) */
                if constexpr (rn_lazy<N, FULL>()) {
                    if (s.lazy)
                        mul_rn<N, true>(s.wim2, s.wim, s.wim, 0);
                    s.lazy = false;
                }
                uint32_t* st = tp->state + (size_t)local * 4 * MDZG_VS(N);
                store_mp<N>(st, s.wre);
                store_mp<N>(st + MDZG_VS(N), s.wim);
                store_mp<N>(st + 2 * MDZG_VS(N), s.wre2);
                store_mp<N>(st + 3 * MDZG_VS(N), s.wim2);
                tp->wzs[local] = wz;
                active = false;
            }
        }
    }
}

#endif
