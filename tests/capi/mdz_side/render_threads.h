/* Test-side declaration of the render-thread stop query the shim calls
   (the pool in tests/capi/shim_pool.c defines it).  Not MDZ's header. */
#ifndef MDZ_SIDE_RENDER_THREADS_H
#define MDZ_SIDE_RENDER_THREADS_H

typedef struct rthdata rthdata;
int rth_render_should_stop(rthdata* rth);

#endif
