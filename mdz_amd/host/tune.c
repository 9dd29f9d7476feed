/*
 * tune.c -- options, the tuned table, and the autotuner (C99).
 *
 * The reference's only tuning knob is the number of render threads
 * (src/main.c:55-59, :243-273).  Here each kernel class has two: the
 * resident blocks per CU its persistent grid is sized for, and the
 * iterations a lane runs between queue refills (chunk).  mdzg_autotune
 * measures a probe tile of the class at every candidate pair and keeps the
 * fastest; mdzg_tune_save / mdzg_tune_load write and read the table, and
 * mdzg_create loads the table shipped beside the library.
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mdzg_host.h"

int mdzg_lib_dir(char* buf, size_t cap)
{
    Dl_info info;
    if (!buf || cap < 2 || !dladdr((void*)&mdzg_lib_dir, &info) || !info.dli_fname)
        return -1;
    snprintf(buf, cap, "%s", info.dli_fname);
    char* slash = strrchr(buf, '/');
    if (!slash)
        snprintf(buf, cap, ".");
    else
        *slash = '\0';
    return 0;
}

int mdzg_option_set(mdzg_ctx* c, int option, long long v)
{
    if (!c)
        return MDZG_ERR_ARG;
    switch (option) {
    case MDZG_OPT_GRAPHS:
        if (v != 0 && v != 1)
            return MDZG_ERR_ARG;
        c->graphs = (int)v;
        return MDZG_OK;
    case MDZG_OPT_BAND_PIXELS:
        if (v < 1 || v > (1ll << 30))
            return MDZG_ERR_ARG;
        c->band_pixels = v;
        return MDZG_OK;
    case MDZG_OPT_BAND_INFLIGHT:
        if (v < 1 || v > 16)
            return MDZG_ERR_ARG;
        c->band_inflight = (int)v;
        return MDZG_OK;
    default:
        return MDZG_ERR_ARG;
    }
}

int mdzg_option_get(mdzg_ctx* c, int option, long long* v)
{
    if (!c || !v)
        return MDZG_ERR_ARG;
    switch (option) {
    case MDZG_OPT_GRAPHS: *v = c->graphs; return MDZG_OK;
    case MDZG_OPT_BAND_PIXELS: *v = c->band_pixels; return MDZG_OK;
    case MDZG_OPT_BAND_INFLIGHT: *v = c->band_inflight; return MDZG_OK;
    default: return MDZG_ERR_ARG;
    }
}

/* Table format, one entry a line ('#' starts a comment):
     kernel <id> <blocks_per_cu> <chunk>
     option <name> <value>         name: graphs, band_pixels, band_inflight */
int mdzg_tune_save(mdzg_ctx* c, const char* path)
{
    if (!c || !path)
        return MDZG_ERR_ARG;
    FILE* f = fopen(path, "w");
    if (!f)
        return MDZG_ERR_ARG;
    fprintf(f, "# mdzg tuned table: kernel <id> <blocks_per_cu> <chunk>; option <name> <value>\n");
    for (int k = 0; k < mdzg_dev_nkernels(); ++k)
        fprintf(f, "kernel %d %d %d\n", k, c->occ[k], c->chunk[k]);
    fprintf(f, "option graphs %d\n", c->graphs);
    fprintf(f, "option band_pixels %lld\n", c->band_pixels);
    fprintf(f, "option band_inflight %d\n", c->band_inflight);
    return fclose(f) == 0 ? MDZG_OK : MDZG_ERR_ARG;
}

int mdzg_tune_load(mdzg_ctx* c, const char* path)
{
    if (!c || !path)
        return MDZG_ERR_ARG;
    FILE* f = fopen(path, "r");
    if (!f)
        return MDZG_ERR_ARG;
    char line[256];
    int rc = MDZG_OK;
    while (fgets(line, sizeof line, f)) {
        if (line[0] == '#' || line[0] == '\n')
            continue;
        int id, bpc, chunk;
        char name[64];
        long long v;
        if (sscanf(line, "kernel %d %d %d", &id, &bpc, &chunk) == 3) {
            if (mdzg_tune_set(c, id, bpc, chunk) != MDZG_OK)
                rc = MDZG_ERR_PARSE;
        } else if (sscanf(line, "option %63s %lld", name, &v) == 2) {
            int opt = !strcmp(name, "graphs") ? MDZG_OPT_GRAPHS
                      : !strcmp(name, "band_pixels") ? MDZG_OPT_BAND_PIXELS
                      : !strcmp(name, "band_inflight") ? MDZG_OPT_BAND_INFLIGHT : -1;
            if (opt < 0 || mdzg_option_set(c, opt, v) != MDZG_OK)
                rc = MDZG_ERR_PARSE;
        } else {
            rc = MDZG_ERR_PARSE;
        }
    }
    fclose(f);
    return rc;
}

/* kernel time of one render of a probe at the context's current pair: the
   fastest of reps */
static int probe_ms(mdzg_ctx* c, const mdzg_tile* probe, int reps, double* ms)
{
    double best = 1e300;
    for (int r = 0; r < reps; ++r) {
        mdzg_stats st;
        int rc = mdzg_render_batch(c, probe, 1, MDZG_BATCH_PERSISTENT, &st);
        if (rc != MDZG_OK)
            return rc;
        if (st.kernel_ms < best)
            best = st.kernel_ms;
    }
    *ms = best;
    return MDZG_OK;
}

#define MAX_PROBES 8

/* the pair's cost: the mean over the probes of its time over the time of the
   pair in place (t0), so that every probe weighs alike */
static int pair_cost(mdzg_ctx* c, const mdzg_tile* probes, int np, int reps, const double* t0,
                     double* cost, double* ms)
{
    double s = 0.0, tot = 0.0;
    for (int i = 0; i < np; ++i) {
        double t;
        int rc = probe_ms(c, &probes[i], reps, &t);
        if (rc != MDZG_OK)
            return rc;
        s += t0 ? t / t0[i] : 1.0;
        tot += t;
    }
    *cost = s / np;
    *ms = tot;
    return MDZG_OK;
}

int mdzg_autotune(mdzg_ctx* c, const mdzg_tile* probes, int nprobes, int reps,
                  mdzg_tune_result* out)
{
    if (!c || !probes || nprobes < 1 || nprobes > MAX_PROBES || reps < 1 || reps > 100)
        return MDZG_ERR_ARG;
    const int kid = mdzg_kernel_id(probes[0].view.path, probes[0].view.precision);
    if (kid < 0)
        return MDZG_ERR_RANGE;
    for (int i = 1; i < nprobes; ++i)
        if (mdzg_kernel_id(probes[i].view.path, probes[i].view.precision) != kid)
            return MDZG_ERR_ARG;
    const int occ0 = c->occ[kid], chunk0 = c->chunk[kid];
    const int maxb = mdzg_dev_occupancy(kid, 1) > 0 ? mdzg_dev_occupancy(kid, 1) : 1;
    static const int chunks_rn[] = { 1, 2, 4, 8, 16, 32 };
    static const int chunks_fp64[] = { 64, 128, 256, 512, 1024 };
    const int* chunks = kid == 0 ? chunks_fp64 : chunks_rn;
    const int nch = kid == 0 ? 5 : 6;
    /* warm-up (code objects, buffers), then the pair in place on every probe */
    double t0[MAX_PROBES], t0sum = 0.0;
    for (int i = 0; i < nprobes; ++i) {
        int rc = probe_ms(c, &probes[i], 1, &t0[i]);
        if (rc == MDZG_OK)
            rc = probe_ms(c, &probes[i], reps, &t0[i]);
        if (rc != MDZG_OK)
            return rc;
        t0sum += t0[i];
    }
    /* a candidate replaces the pair in place only when its cost is lower by
       more than the run-to-run spread of a probe (2 %) */
    double best = 1e300, best_ms = 0.0;
    int bb = occ0, bc = chunk0, n = 1, rc = MDZG_OK;
    for (int b = 1; b <= maxb && rc == MDZG_OK; ++b)
        for (int q = 0; q < nch && rc == MDZG_OK; ++q) {
            if (b == occ0 && chunks[q] == chunk0)
                continue;
            c->occ[kid] = b;
            c->chunk[kid] = chunks[q];
            double cost, ms;
            rc = pair_cost(c, probes, nprobes, reps, t0, &cost, &ms);
            ++n;
            if (rc == MDZG_OK && cost < best) {
                best = cost;
                best_ms = ms;
                bb = b;
                bc = chunks[q];
            }
        }
    if (!(best < 0.98)) {
        best_ms = t0sum;
        bb = occ0;
        bc = chunk0;
    }
    c->occ[kid] = bb;
    c->chunk[kid] = bc;
    if (out) {
        out->kernel_id = kid;
        out->blocks_per_cu = bb;
        out->chunk = bc;
        out->best_ms = best_ms;
        out->default_ms = t0sum;
        out->candidates = n;
    }
    return rc;
}

int mdzg_rate_probe(mdzg_ctx* c, int kind, int blocks_per_cu, double* ops_per_s)
{
    if (!c || !ops_per_s || (kind != 0 && kind != 1) || blocks_per_cu < 1 || blocks_per_cu > 32)
        return MDZG_ERR_ARG;
    hipSetDevice(c->device);
    const int iters = 4096;
    double ms = 0.0;
    if (mdzg_dev_rate_probe(kind, c->ncu, blocks_per_cu, iters, &ms) != 0 || ms <= 0.0)
        return MDZG_ERR_DEVICE;
    const double lanes = (double)c->ncu * blocks_per_cu * 256.0;
    *ops_per_s = lanes * 4.0 * iters / (ms * 1e-3);
    return MDZG_OK;
}
