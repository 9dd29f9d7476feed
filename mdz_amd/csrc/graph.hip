/*
 * graph.hip -- launch nodes and their replay as cached HIP graphs.
 *
 * A batch or a serving slice is a small launch plan: an optional root
 * kernel (the sentinel fill of the counts, or the reset of the queue
 * counters), the persistent or slice kernels of its arithmetic classes,
 * which run concurrently after the root, and an optional tail that waits
 * for all of them (the colouring pass of an image job).  Launched one by
 * one, such a plan costs a stream fork, one host launch per kernel and a
 * join.  As a HIP graph it is one launch, and the fork and join are graph
 * edges.
 *
 * The cache keeps instantiated graphs by shape: the stream, the number of
 * root, concurrent and tail nodes, and each node's kernel and block size.
 * A plan of a known shape replays its graph after replacing every node's
 * grid and arguments (hipGraphExecKernelNodeSetParams), which costs far less
 * than instantiating a graph; HIP takes a kernel node's arguments when the
 * graph is launched, so a replay in flight keeps its own.  A lock covers the
 * update and the launch, so two threads never interleave them on one entry;
 * graphs of one stream run in that stream's order.  No graph holds a memset
 * node: fills are kernels.
 */
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <mutex>

#include "mdzg_dev.h"

#define GC_ENTRIES 32
#define GC_MAXN (2 + 2 * MDZG_NKERNELS)

namespace {

struct gentry {
    bool used;
    hipStream_t st;
    int nroot, npar, ntail;
    const void* func[GC_MAXN];
    unsigned block[GC_MAXN], shmem[GC_MAXN];
    hipGraph_t g;
    hipGraphExec_t x;
    hipGraphNode_t node[GC_MAXN];
    uint64_t tick;
};

void params_of(const mdzg_gnode& n, hipKernelNodeParams& p, void** argp)
{
    for (int i = 0; i < n.nargs; ++i)
        argp[i] = (void*)((const char*)n.buf + n.off[i]);
    memset(&p, 0, sizeof p);
    p.func = const_cast<void*>(n.func);
    p.gridDim = dim3(n.grid);
    p.blockDim = dim3(n.block);
    p.sharedMemBytes = n.shmem;
    p.kernelParams = argp;
    p.extra = nullptr;
}

} /* namespace */

struct mdzg_gcache {
    std::mutex m;
    gentry e[GC_ENTRIES];
    uint64_t tick;
    uint64_t hits, builds;
};

extern "C" {

mdzg_gcache* mdzg_gcache_create(void)
{
    mdzg_gcache* gc = new (std::nothrow) mdzg_gcache();
    if (gc) {
        memset(gc->e, 0, sizeof gc->e);
        gc->tick = gc->hits = gc->builds = 0;
    }
    return gc;
}

void mdzg_gcache_destroy(mdzg_gcache* gc)
{
    if (!gc)
        return;
    for (int i = 0; i < GC_ENTRIES; ++i)
        if (gc->e[i].used) {
            hipGraphExecDestroy(gc->e[i].x);
            hipGraphDestroy(gc->e[i].g);
        }
    delete gc;
}

void mdzg_gcache_counts(mdzg_gcache* gc, uint64_t* hits, uint64_t* builds)
{
    std::lock_guard<std::mutex> l(gc->m);
    *hits = gc->hits;
    *builds = gc->builds;
}

/* one node, launched at once */
int mdzg_dev_node_launch(const mdzg_gnode* n, hipStream_t st)
{
    void* argp[MDZG_GNODE_ARGS];
    for (int i = 0; i < n->nargs; ++i)
        argp[i] = (void*)((const char*)n->buf + n->off[i]);
    return hipLaunchKernel(n->func, dim3(n->grid), dim3(n->block), argp, n->shmem, st) ==
                   hipSuccess
               ? 0
               : -3;
}

/* Run root (may be null), then par[0..npar) concurrently, then tail (may be
   null), on stream st as one graph launch.  0, or -3 on a runtime error. */
int mdzg_graph_run(mdzg_gcache* gc, const mdzg_gnode* root, const mdzg_gnode* par, int npar,
                   const mdzg_gnode* tail, hipStream_t st)
{
    const int nroot = root ? 1 : 0, ntail = tail ? 1 : 0;
    const int n = nroot + npar + ntail;
    if (!gc || npar < 0 || n < 1 || n > GC_MAXN)
        return -1;
    const mdzg_gnode* nodes[GC_MAXN];
    int k = 0;
    if (root)
        nodes[k++] = root;
    for (int i = 0; i < npar; ++i)
        nodes[k++] = &par[i];
    if (tail)
        nodes[k++] = tail;

    std::lock_guard<std::mutex> l(gc->m);
    gentry* hit = nullptr;
    for (int i = 0; i < GC_ENTRIES && !hit; ++i) {
        gentry& e = gc->e[i];
        if (!e.used || e.st != st || e.nroot != nroot || e.npar != npar || e.ntail != ntail)
            continue;
        bool same = true;
        for (int j = 0; j < n && same; ++j)
            same = e.func[j] == nodes[j]->func && e.block[j] == nodes[j]->block &&
                   e.shmem[j] == nodes[j]->shmem;
        if (same)
            hit = &e;
    }
    void* argp[GC_MAXN][MDZG_GNODE_ARGS];
    hipKernelNodeParams p[GC_MAXN];
    for (int j = 0; j < n; ++j)
        params_of(*nodes[j], p[j], argp[j]);
    if (hit) {
        for (int j = 0; j < n; ++j)
            if (hipGraphExecKernelNodeSetParams(hit->x, hit->node[j], &p[j]) != hipSuccess)
                return -3;
        gc->hits++;
    } else {
        /* a new shape: build, instantiate, and take the least recently used slot */
        gentry* slot = &gc->e[0];
        for (int i = 0; i < GC_ENTRIES; ++i) {
            if (!gc->e[i].used) {
                slot = &gc->e[i];
                break;
            }
            if (gc->e[i].tick < slot->tick)
                slot = &gc->e[i];
        }
        if (slot->used) {
            hipGraphExecDestroy(slot->x);
            hipGraphDestroy(slot->g);
            slot->used = false;
        }
        hipGraph_t g;
        if (hipGraphCreate(&g, 0) != hipSuccess)
            return -3;
        hipGraphNode_t nd[GC_MAXN];
        bool ok = true;
        int j = 0;
        if (root) {
            ok = hipGraphAddKernelNode(&nd[0], g, nullptr, 0, &p[0]) == hipSuccess;
            j = 1;
        }
        for (int i = 0; i < npar && ok; ++i, ++j)
            ok = hipGraphAddKernelNode(&nd[j], g, root ? &nd[0] : nullptr, root ? 1 : 0, &p[j]) ==
                 hipSuccess;
        if (tail && ok) {
            if (npar > 0)
                ok = hipGraphAddKernelNode(&nd[j], g, &nd[nroot], (size_t)npar, &p[j]) ==
                     hipSuccess;
            else
                ok = hipGraphAddKernelNode(&nd[j], g, nroot ? &nd[0] : nullptr, (size_t)nroot,
                                           &p[j]) == hipSuccess;
            ++j;
        }
        hipGraphExec_t x;
        if (!ok || hipGraphInstantiate(&x, g, nullptr, nullptr, 0) != hipSuccess) {
            hipGraphDestroy(g);
            return -3;
        }
        slot->used = true;
        slot->st = st;
        slot->nroot = nroot;
        slot->npar = npar;
        slot->ntail = ntail;
        for (int q = 0; q < n; ++q) {
            slot->func[q] = nodes[q]->func;
            slot->block[q] = nodes[q]->block;
            slot->shmem[q] = nodes[q]->shmem;
            slot->node[q] = nd[q];
        }
        slot->g = g;
        slot->x = x;
        hit = slot;
        gc->builds++;
    }
    hit->tick = ++gc->tick;
    return hipGraphLaunch(hit->x, st) == hipSuccess ? 0 : -3;
}

} /* extern "C" */
