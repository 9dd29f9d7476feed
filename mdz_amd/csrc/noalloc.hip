/*
 * noalloc.hip -- the library's static device scratch for the no-allocation
 * entry point (mdzg_render_tile_noalloc, runtime.c).
 *
 * A render needs device memory beyond its counts: the tile descriptor, the
 * queue counter, the cancel flag and the per-column / per-row start values.
 * mdzg_render_tile takes them from a buffer the context allocates; the
 * no-allocation variant takes them from this file-scope array instead, so a
 * caller that cannot allocate device memory only brings its counts buffer.
 *
 * In host code `scratch` names a host-side stand-in, not device memory: the
 * only host use of the name is the hipGetSymbolAddress call below, made once
 * per device with that device current.  Everything else (the copy of the
 * staged descriptor and start values, the kernel arguments) uses the device
 * address it returns, kept per device.  Each GPU holds its own copy of the
 * array until the process ends; every call on a GPU shares it.
 */
#include <hip/hip_runtime.h>
#include <stddef.h>

#include <mutex>

#include "../../include/mdzg.h"
#include "mdzg_dev.h"

static __device__ __align__(256) unsigned char scratch[MDZG_NOALLOC_SCRATCH_BYTES];

#define MAX_DEVICES 64
static void* g_scratch_ptr[MAX_DEVICES];
static std::mutex g_scratch_lock;

extern "C" {

/* Device address of this device's scratch array (the current device must be
   `device`).  0 on success, -3 when the runtime cannot resolve the symbol. */
int mdzg_dev_scratch(int device, void** ptr, size_t* bytes)
{
    if (device < 0 || device >= MAX_DEVICES || !ptr)
        return -1;
    std::lock_guard<std::mutex> g(g_scratch_lock);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return -3;
        }
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    if (bytes)
        *bytes = MDZG_NOALLOC_SCRATCH_BYTES;
    return 0;
}

} /* extern "C" */
