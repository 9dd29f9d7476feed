/* k_probe.hip -- the per-operation probe kernels of the RN arithmetic
 * (k_rn_op_probe in k_escape.h) for the 128- and 1024-bit legs, in a code
 * object of their own so that the render kernels' objects hold only the
 * render kernels.  Exports mdzg_kprobe_rn4() and mdzg_kprobe_rn32(). */
#include "k_escape.h"

extern "C" const void* mdzg_kprobe_rn4(void)
{
    return (const void*)k_rn_op_probe<4>;
}

extern "C" const void* mdzg_kprobe_rn32(void)
{
    return (const void*)k_rn_op_probe<32>;
}
