/* k_rn.hip -- the RN<N> kernels of k_escape.h for one word count, selected by
 * -DMDZG_N=<N> at compile time (one object per N, built in parallel).
 * Exports mdzg_kget_rn<N>(full, &tile, &persist, &slice) to the launch table. */
#include "k_escape.h"

#ifndef MDZG_N
#error "build with -DMDZG_N=<words>"
#endif

#define KGET_NAME2(n) mdzg_kget_rn##n
#define KGET_NAME(n) KGET_NAME2(n)

extern "C" void KGET_NAME(MDZG_N)(int full, const void** tile, const void** persist,
                                  const void** slice)
{
    if (full) {
        *tile = (const void*)k_rn_tile<MDZG_N, true>;
        *persist = (const void*)k_rn_persist<MDZG_N, true>;
        *slice = (const void*)k_rn_slice<MDZG_N, true>;
    } else {
        *tile = (const void*)k_rn_tile<MDZG_N, false>;
        *persist = (const void*)k_rn_persist<MDZG_N, false>;
        *slice = (const void*)k_rn_slice<MDZG_N, false>;
    }
}

