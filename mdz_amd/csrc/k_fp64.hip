/* k_fp64.hip -- the FP64 kernels of k_escape.h, exported to the launch table. */
#define MDZG_FP64_KERNELS 1
#include "k_escape.h"

extern "C" void mdzg_kget_fp64(const void** tile, const void** persist, const void** slice)
{
    *tile = (const void*)k_fp64_tile;
    *persist = (const void*)k_fp64_persist;
    *slice = (const void*)k_fp64_slice;
}
