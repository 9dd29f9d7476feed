"""Benchmark: G pixel-iterations/s of the escape-time path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config 2|3|4|5|gmp]
                    [--no-cpu-baseline]

A step is one pass of the hot path over one batch of synthetic input.  The
default workload is config 2 (the configuration the metric is quoted on):
Mandelbrot on the MPFR path at p = 128, 640x480 at the seahorse-valley centre,
width 1e-25, depth 5000; every pixel is inside, so a step is exactly
640*480*5000 = 1.536e9 pixel-iterations.  The line also carries the FP64
figure (config 1, 240x180, depth 1000), the 1024-bit figure (a band of
config 5 rows) and, with --config 3/4/5/gmp, the batched and serving forms,
the full deep render and the GMP path.

pixel-iterations = sum over samples of (count if count > 0 else depth): the
iterations the reference loop executes.  Algorithmic operations per
pixel-iteration: 7 FP64 operations (double path), 2n^2+n 32x32-bit
multiply-accumulates with n = p/32 (RN paths).  The roofline peak is MI355X's
78.6 TFLOP/s FP64 vector rate; a 32x32->64 multiply-accumulate
(v_mad_u64_u32) issues at the same rate as an FP64 FMA on gfx950 and is
counted as 2 operations like the FMA.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "G pixel-iter/s at FP64/128/1024-bit vs MPFR CPU; p50/p99 tile latency"
PEAK_TOPS = 78.6
PEAK_HBM_GBS = 8000.0
# HBM bytes per launch of k_colour on the 3840x2160 colouring workload by aa:
# rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1 (scaled by a 256 MiB copy's readings)
COLOUR_TRAFFIC_BYTES = {1: 66.4e6, 2: 166.0e6}
# rocprofv3 counter figures of k_rn_persist<4,true> on config 2 (HBM bytes per
# launch from FETCH_SIZE / WRITE_SIZE scaled by a plain copy's readings, VALU
# instructions per launch from SQ_INSTS_VALU), read at run time from the
# profile summary they were taken in, which names the command and the build
CONFIG2_COUNTERS = os.path.join(ROOT, "profiles", "k_rn_persist4_config2_counters.md")


def profiled_counters(path):
    """The `bench-counters: {json}` line of a profile summary, with the file
    it came from, or None when the file or the line is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("bench-counters:"):
                    d = json.loads(line.split(":", 1)[1])
                    d["source"] = os.path.relpath(path, ROOT)
                    return d
    except OSError:
        pass
    return None

# VALU issue limit: 256 CUs x one wave64 instruction per clock x 2.4 GHz peak clock
PEAK_VALU_ISSUE = 256 * 2.4e9
# VALU instructions per wave-iteration of the FP64 kernel's run loop: the
# formula's 7 FP64 operations and the exact doubling, one compare, one add
FP64_VALU_PER_ITER = 10


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def counted_cores():
    """Core count as MDZ reads it (src/main.c:243-273: cpuN lines of /proc/stat)."""
    n = 0
    try:
        with open("/proc/stat") as f:
            for line in f:
                if line.startswith("cpu") and line[3:4].isdigit():
                    n += 1
    except OSError:
        pass
    return n or (os.cpu_count() or 1)


def host_cores():
    """Threads for the CPU baselines: the counted cores, capped by the job's
    share of a host that serves several GPUs (OMP_NUM_THREADS)."""
    n = counted_cores()
    cap = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return min(n, cap) if cap > 0 else n


# how MDZ's loops are built for the CPU baselines (oracle/Makefile)
REF_FLAGS = {"O0": "gcc -Wall -Wextra -std=gnu99 -ggdb (src/Makefile:11,17, the reference default)",
             "O2": "gcc -Wall -Wextra -std=gnu99 -O2 -ffp-contract=off"}


def cpu_baseline(rect, precision, depth, W, H, line_step, threads, path=0, build="O0",
                 lines=None, cols=None, sample_name="config 2", repeat=1):
    """MDZ's loop (MPFR, GMP or long double: src/frac_*.c compiled by
    oracle/Makefile with the flags of `build`) or the double restatement
    (path 2, the oracle's -O2 C loop) on the host cores, timed on a sample:
    every line_step-th row, or the lines x cols crop; repeat times."""
    from oracle import oracle as O
    kind = "reference"
    if path == O.PATH_DOUBLE:
        kind = "port"
        O.use_reference(False)
    else:
        try:
            O.use_reference(True, build)
        except Exception:
            kind = "port"
            O.use_reference(False)
    if lines is None:
        nl = (H + line_step - 1) // line_step
        lines = (0, nl, line_step)
    el, pi = 0.0, 0.0
    for _ in range(repeat):
        counts, e = O.render(rect, precision=precision, depth=depth, width=W, height=H,
                             lines=lines, cols=cols, threads=threads, timed=True, path=path)
        el += e
        pi += float(((counts > 0) * counts + (counts == 0) * depth).sum())
    O.use_reference(False)
    where = (f"rows {lines[0]},{lines[0] + lines[2]},.. ({lines[1]} rows)" if cols is None
             else f"{lines[1]}x{cols[1]} crop at row {lines[0]} col {cols[0]}")
    return dict(value=pi / el / 1e9, unit="G pixel-iter/s", cores=threads, kind=kind,
                build=(REF_FLAGS[build] if kind == "reference" else
                       "oracle restatement, gcc -O2 -ffp-contract=off"),
                counted_cores=counted_cores(),
                sample=f"{sample_name} {where} x {W if cols is None else cols[1]} px, "
                       f"{pi:.3e} pixel-iterations{' x' + str(repeat) if repeat > 1 else ''} "
                       f"in {el:.2f} s, {threads} threads claiming rows")


def config3_classes(ctx, M, WL, np, specs):
    """Config 3 by arithmetic class: each class's tiles alone, as one
    persistent launch and as one launch per tile.  The persistent batch of
    all tiles cannot take less than the sum of the classes' persistent times
    less what running the classes concurrently overlaps, and the per-tile
    form takes the sum of its classes' per-tile times; their ratio bounds
    what batching gains on this tile distribution."""
    by = {}
    for sp in specs:
        by.setdefault(sp["precision"], []).append(sp)
    out = {}
    for p in sorted(by):
        tiles = [(WL.spec_view(sp), None, None) for sp in by[p]]
        _, sp_ = ctx.render_batch(tiles, M.BATCH_PERSISTENT)
        _, st_ = ctx.render_batch(tiles, M.BATCH_PER_TILE)
        npix = sum(sp["side"] ** 2 for sp in by[p])
        out[str(p)] = dict(tiles=len(by[p]), pixels=npix, pixel_iters=sp_["pixel_iters"],
                           persistent_ms=sp_["kernel_ms"], per_tile_ms=st_["kernel_ms"],
                           persistent_gpix_per_s=sp_["pixel_iters"] / sp_["kernel_ms"] / 1e6,
                           per_tile_gpix_per_s=st_["pixel_iters"] / st_["kernel_ms"] / 1e6,
                           per_tile_over_persistent=st_["kernel_ms"] / sp_["kernel_ms"])
        log(f"config 3 class p={p}: {len(by[p])} tiles, persistent {sp_['kernel_ms']:.0f} ms, "
            f"per-tile {st_['kernel_ms']:.0f} ms")
    sp_sum = sum(v["persistent_ms"] for v in out.values())
    st_sum = sum(v["per_tile_ms"] for v in out.values())
    return dict(classes=out, sum_class_persistent_ms=sp_sum, sum_class_per_tile_ms=st_sum,
                bound_per_tile_over_persistent=st_sum / sp_sum)


def config3_batched(ctx, M, WL, np, ntiles=256, classes=False, graphs_off=False):
    """Config 3: the seeded tiles in one persistent launch per arithmetic class
    (concurrent classes) against one plain launch per tile; counts must match.
    classes: also each class alone (config3_classes); graphs_off: also the
    persistent batch with its launches made one by one."""
    specs = WL.config3_specs(ntiles)
    t0 = time.perf_counter()
    tiles = [(WL.spec_view(sp), None, None) for sp in specs]
    log(f"config 3: {len(tiles)} tile views in {time.perf_counter() - t0:.2f} s")
    outs_p, st_p = ctx.render_batch(tiles, M.BATCH_PERSISTENT)
    log(f"config 3 persistent: {st_p['kernel_ms']:.0f} ms kernel, {st_p['wall_ms']:.0f} ms wall")
    outs_t, st_t = ctx.render_batch(tiles, M.BATCH_PER_TILE)
    log(f"config 3 per-tile:   {st_t['kernel_ms']:.0f} ms kernel, {st_t['wall_ms']:.0f} ms wall")
    same = all(np.array_equal(x, y) for x, y in zip(outs_p, outs_t))
    by = {}
    for sp in specs:
        by[sp["precision"]] = by.get(sp["precision"], 0) + 1
    gp = st_p["pixel_iters"] / (st_p["wall_ms"] * 1e-3) / 1e9
    gt = st_t["pixel_iters"] / (st_t["wall_ms"] * 1e-3) / 1e9
    extra = {}
    if graphs_off:
        ctx.option_set(M.OPT_GRAPHS, 0)
        _, st_e = ctx.render_batch(tiles, M.BATCH_PERSISTENT)
        ctx.option_set(M.OPT_GRAPHS, 1)
        extra["persistent_without_graph"] = dict(kernel_ms=st_e["kernel_ms"],
                                                 wall_ms=st_e["wall_ms"])
        log(f"config 3 persistent, launches one by one: {st_e['wall_ms']:.0f} ms wall")
    if classes:
        extra["by_class"] = config3_classes(ctx, M, WL, np, specs)
    return dict(
        metric=METRIC, value=gp, steps=1, warmup=0, ms_per_step=st_p["wall_ms"],
        dtype="u32 limbs RN p in {64..1024} + f64",
        config=dict(workload=f"config3 {len(specs)} seeded tiles (seed 0x4D445A), sides "
                    "64-512, log10 width U[-250,-2], depth log-U[1e3,5e4], 60/20/20 "
                    "Mandelbrot/Burning Ship/Celtic", precision_classes=by,
                    pixel_iters=st_p["pixel_iters"]),
        batched=dict(persistent_gpix_per_s=gp, per_tile_gpix_per_s=gt, speedup=gp / gt,
                     persistent_kernel_ms=st_p["kernel_ms"], per_tile_kernel_ms=st_t["kernel_ms"],
                     launches_persistent=st_p["launches"], launches_per_tile=st_t["launches"],
                     identical_counts=bool(same), **extra))


def gmp_path(ctx, M, WL, a, cpu=True):
    """MDZ's GMP path (truncating mpf, src/frac_mandel.c:55-80) on the config 2
    view: full frames at p = 128 (4-limb values)."""
    rect = M.rect_from_center(WL.CONFIG2_CENTRE[0], WL.CONFIG2_CENTRE[1], "1e-25", 128, 640,
                              480, coords=True)
    v = M.make_view(rect, precision=128, depth=5000, width=640, height=480, path=M.PATH_GMP)
    ctx.render(v, lines=(0, 8), cols=(0, 640))
    steps = max(1, min(a.steps, 3))
    t0 = time.perf_counter()
    kms, pis = [], []
    for _ in range(steps):
        _, st = ctx.render(v)
        kms.append(st["kernel_ms"])
        pis.append(st["pixel_iters"])
    el = time.perf_counter() - t0
    r = dict(value=sum(pis) / el / 1e9, steps=steps, warmup=1, ms_per_step=1e3 * el / steps,
             dtype="u64 limbs, GMP mpf PREC 3 (p=128)",
             config=dict(workload="config2 view on the GMP path (truncating mpf, p=128) "
                         "640x480 depth 5000", kernel="k_gmp_persist<4>"),
             kernel_gpix_per_s=pis[0] / (sum(kms) / len(kms) * 1e-3) / 1e9)
    log(f"gmp path: {r['kernel_gpix_per_s']:.2f} G pixel-iter/s (kernel)")
    # rows of the same view at the other storage buckets' precisions, the
    # MPFR path on the same rows for scale (96 rows at p <= 256, 32 above)
    bands = {}
    for p in (256, 512, 1024):
        rows = 96 if p <= 256 else 32
        out = {}
        for name, path in (("gmp", M.PATH_GMP), ("mpfr", M.PATH_MPFR)):
            rp = M.rect_from_center(WL.CONFIG2_CENTRE[0], WL.CONFIG2_CENTRE[1], "1e-25", p, 640,
                                    480, coords=(path == M.PATH_GMP))
            vp = M.make_view(rp, precision=p, depth=5000, width=640, height=480, path=path)
            ctx.render(vp, lines=(0, 4))
            _, st = ctx.render(vp, lines=(0, rows))
            out[name] = st["pixel_iters"] / (st["kernel_ms"] * 1e-3) / 1e9
        bands[f"p{p}"] = dict(rows=rows, gmp_gpix_per_s=out["gmp"], mpfr_gpix_per_s=out["mpfr"])
        log(f"gmp p={p}: {out['gmp']:.2f} G pixel-iter/s, mpfr path {out['mpfr']:.2f}")
    r["bands"] = bands
    if cpu:
        r["cpu_baseline"] = cpu_baseline(list(rect), 128, 5000, 640, 480, 16, host_cores(),
                                         path=3)
        r["cpu_baseline"]["sample"] = r["cpu_baseline"]["sample"].replace("config 2", "config 2 GMP")
        log(f"gmp cpu baseline: {r['cpu_baseline']['value']:.4f} G pixel-iter/s")
    return r


def colour_pass(ctx, M, np, a, cpu=True):
    """The colouring pass (k_colour: palette_apply / do_anti_aliasing) on a
    3840x2160 image, one sample per pixel and 2x2 anti-aliasing, synthetic
    counts (0..20000, a tenth never escaped), a 256-colour interpolated
    palette.  HBM-bound: 4 aa^2 bytes read + 4 written per output pixel."""
    W, H = 3840, 2160
    rng = np.random.default_rng(7)
    pal = M.make_palette([int(x) for x in rng.integers(0, 1 << 24, 256)], offset=148,
                         colour_scale=0.02539073500000000108, interpolate=True)
    out = {}
    steps = max(3, min(a.steps, 20))
    for aa in (1, 2):
        c = rng.integers(1, 20000, (H * aa, W * aa), dtype=np.int64).astype(np.uint32)
        c[rng.random(c.shape) < 0.1] = 0
        for _ in range(2):
            ctx.colour(c, aa, pal)
        kms, walls = [], []
        for _ in range(steps):
            _, st = ctx.colour(c, aa, pal, stats=True)
            kms.append(st["kernel_ms"])
            walls.append(st["wall_ms"])
        kavg = sum(kms) / len(kms)
        nbytes = 4.0 * aa * aa * W * H + 4.0 * W * H
        gbs = nbytes / (kavg * 1e-3) / 1e9
        out[f"aa{aa}"] = dict(kernel_ms=kavg, wall_ms_p50=sorted(walls)[len(walls) // 2],
                              mpix_per_s=W * H / (kavg * 1e-3) / 1e6,
                              roofline=dict(bound="hbm", achieved=gbs, peak=PEAK_HBM_GBS,
                                            unit="GB/s", frac=gbs / PEAK_HBM_GBS,
                                            traffic=COLOUR_TRAFFIC_BYTES[aa],
                                            bytes_per_launch=nbytes))
        log(f"colour aa={aa}: {kavg:.3f} ms, {gbs:.0f} GB/s")
        if cpu and aa == 2:
            from oracle import oracle as O
            rows = 64
            t0 = time.perf_counter()
            O.colour(c[:rows * aa], aa, M.palette_colours(pal), offset=148,
                     colour_scale=0.02539073500000000108, interpolate=True)
            el = time.perf_counter() - t0
            out["cpu_baseline"] = dict(value=rows * W / el / 1e6, unit="M output pixels/s",
                                       cores=1, kind="port",
                                       sample=f"{rows} rows x {W} pixels of the aa=2 image, "
                                              "MDZ's colouring restated in C (oracle)")
    return out


def crop_tiles(WL, n, seed=0x5E):
    """64x64 crops of the config 2 view at seeded positions (p=128, depth 5000)."""
    import random
    v = WL.config2_view()
    rng = random.Random(seed)
    return [(v, (rng.randrange(0, 480 - 64), 64), (rng.randrange(0, 640 - 64), 64))
            for _ in range(n)], v


def serve_run(ctx, M, np, mode, tiles, rate, seed=1, palette=None):
    """Poisson arrivals at `rate` requests/s (inf: all at once); latency per request
    is done - submit on the host clock; throughput over first submit .. last done.
    With a palette the requests are image requests (coloured on the device)."""
    import random
    rng = random.Random(seed)
    srv = M.Server(ctx, mode=mode)
    t0 = time.perf_counter()
    tnext = 0.0
    ids = []
    for v, l, c in tiles:
        if rate != float("inf"):
            tnext += rng.expovariate(rate)
            dt = tnext - (time.perf_counter() - t0)
            if dt > 0:
                time.sleep(dt)
        ids.append(srv.submit(v, l, c) if palette is None else
                   srv.submit_image(v, 1, palette, l, c))
    res = [srv.wait(i)[1] for i in ids]
    st = srv.stats()
    srv.close()
    lat = np.array([r["latency_ms"] for r in res])
    span = (max(r["done_ms"] for r in res) - min(r["submit_ms"] for r in res)) * 1e-3
    pix = sum(r["pixel_iters"] for r in res)
    return dict(rate=rate, n=len(tiles), p50_ms=float(np.percentile(lat, 50)),
                p99_ms=float(np.percentile(lat, 99)), mean_ms=float(lat.mean()),
                tiles_per_s=len(tiles) / span, gpix_per_s=pix / span / 1e9,
                slices=st["slices"], hbm_peak=st["hbm_peak"], latencies_ms=lat.tolist())


def _brief(r):
    """a serve_run result without its per-request latencies"""
    return {k: v for k, v in r.items() if k != "latencies_ms"}


def saturation_run(ctx, M, np, mode, tiles):
    """Saturation throughput at steady state: one untimed run of the whole
    request set, then the timed one.  A saturation run that follows a mostly
    idle phase (the unloaded latencies, one request at a time) or a small
    warm-up ran 10-15 % slower on 300 config 2 crops than one that follows
    another busy run (86-89 against 95-99 G pixel-iter/s, graphs on or
    off)."""
    serve_run(ctx, M, np, mode, tiles, float("inf"))
    return serve_run(ctx, M, np, mode, tiles, float("inf"))


def unloaded_latency(ctx, M, np, mode, tiles, reps=15):
    srv = M.Server(ctx, mode=mode)
    lat = []
    for v, l, c in tiles[:reps]:
        i = srv.submit(v, l, c)
        lat.append(srv.wait(i)[1]["latency_ms"])
    srv.close()
    return float(np.median(lat))


def config4_serving(ctx, M, WL, np, fracs=(0.25, 0.5, 0.75, 0.9), n=300):
    """Config 4: Poisson tile requests swept towards saturation, continuous
    batching against one launch per request."""
    tiles, _ = crop_tiles(WL, n)
    out = {}
    serve_run(ctx, M, np, M.SERVE_BATCHED, tiles[:60], float("inf"))     # warm-up
    for name, mode in (("batched", M.SERVE_BATCHED), ("per_request", M.SERVE_PER_REQUEST)):
        unl = unloaded_latency(ctx, M, np, mode, tiles)
        sat = saturation_run(ctx, M, np, mode, tiles)
        log(f"serve {name}: unloaded {unl:.2f} ms, saturation {sat['tiles_per_s']:.0f} tiles/s "
            f"{sat['gpix_per_s']:.1f} G pixel-iter/s")
        sweep = []
        for f in fracs:
            r = serve_run(ctx, M, np, mode, tiles, f * sat["tiles_per_s"])
            r["load"] = f
            sweep.append(r)
            log(f"  load {f:.2f}: {r['tiles_per_s']:.0f} tiles/s p50 {r['p50_ms']:.2f} "
                f"p99 {r['p99_ms']:.2f} ms")
        out[name] = dict(unloaded_ms=unl, saturation=_brief(sat), sweep=[_brief(x) for x in sweep])
    # the same requests as image requests: counts coloured in HBM, colours returned
    rng = np.random.default_rng(7)
    pal = M.make_palette([int(x) for x in rng.integers(0, 1 << 24, 256)], offset=148,
                         colour_scale=0.02539073500000000108, interpolate=True)
    img = serve_run(ctx, M, np, M.SERVE_BATCHED, tiles, float("inf"), palette=pal)
    log(f"serve batched, image requests: saturation {img['tiles_per_s']:.0f} tiles/s "
        f"{img['gpix_per_s']:.1f} G pixel-iter/s")
    out["batched_image_saturation"] = _brief(img)
    # the batched form with each slice's launches made one by one (no graphs)
    ctx.option_set(M.OPT_GRAPHS, 0)
    unl0 = unloaded_latency(ctx, M, np, M.SERVE_BATCHED, tiles)
    sat0 = saturation_run(ctx, M, np, M.SERVE_BATCHED, tiles)
    half0 = serve_run(ctx, M, np, M.SERVE_BATCHED, tiles, 0.5 * sat0["tiles_per_s"])
    ctx.option_set(M.OPT_GRAPHS, 1)
    out["batched_without_graphs"] = dict(unloaded_ms=unl0, saturation=_brief(sat0),
                                         half_load=_brief(half0))
    log(f"serve batched without graphs: unloaded {unl0:.2f} ms, saturation "
        f"{sat0['gpix_per_s']:.1f} G pixel-iter/s, half load p99 {half0['p99_ms']:.2f} ms")
    return out


def config4_mix(ctx, M, WL, np, n=48, fracs=(0.25, 0.5, 0.75), seed=0x5E4):
    """Config 4 on tiles drawn from config 3's distribution (sides 64-512,
    widths 1e-2..1e-250 and so FP64 to 1024-bit tiles, depths 1e3-5e4, three
    formulas): each request's latency is also taken over its own unloaded
    latency (the tile alone on the server), since the tiles' costs span four
    orders of magnitude."""
    specs = WL.config3_specs(n, seed=seed)
    tiles = [(WL.spec_view(sp), None, None) for sp in specs]
    serve_run(ctx, M, np, M.SERVE_BATCHED, tiles[:8], float("inf"))      # warm-up
    srv = M.Server(ctx, mode=M.SERVE_BATCHED)
    alone = []
    for v, l, c in tiles:
        alone.append(srv.wait(srv.submit(v, l, c))[1]["latency_ms"])
    srv.close()
    alone = np.array(alone)

    def slow(r):
        x = np.array(r["latencies_ms"]) / alone
        return dict(p50_slowdown=float(np.percentile(x, 50)),
                    p99_slowdown=float(np.percentile(x, 99)))

    out = dict(workload=f"{n} tiles of config 3's distribution (seed {seed:#x}): "
                        f"{sum(sp['side'] ** 2 for sp in specs)} pixels, precisions "
                        f"{sorted(set(sp['precision'] for sp in specs))}",
               unloaded_ms=dict(p50=float(np.percentile(alone, 50)),
                                p99=float(np.percentile(alone, 99)), sum=float(alone.sum())))
    for name, mode in (("batched", M.SERVE_BATCHED), ("per_request", M.SERVE_PER_REQUEST)):
        sat = serve_run(ctx, M, np, mode, tiles, float("inf"))
        out[name] = dict(saturation=_brief(sat), **slow(sat))
        log(f"mix {name}: saturation {sat['tiles_per_s']:.1f} tiles/s "
            f"{sat['gpix_per_s']:.1f} G pixel-iter/s")
    for name, mode in (("batched", M.SERVE_BATCHED), ("per_request", M.SERVE_PER_REQUEST)):
        sweep = []
        for f in fracs:
            # the same arrival rates for both forms: fractions of the batched saturation
            r = serve_run(ctx, M, np, mode, tiles, f * out["batched"]["saturation"]["tiles_per_s"])
            row = _brief(r)
            row.update(load=f, **slow(r))
            sweep.append(row)
            log(f"  mix {name} load {f:.2f}: {r['gpix_per_s']:.1f} G, p50 {r['p50_ms']:.1f} p99 "
                f"{r['p99_ms']:.1f} ms, p50/p99 slowdown {row['p50_slowdown']:.2f}/"
                f"{row['p99_slowdown']:.2f}")
        out[name]["sweep"] = sweep
    sweep = out["batched"]["sweep"]
    ok = [x for x in sweep if x["p99_slowdown"] <= 2.0]
    best = max(ok, key=lambda x: x["gpix_per_s"]) if ok else None
    pr = out["per_request"]["saturation"]["gpix_per_s"]
    out["batching_target"] = dict(load=best["load"] if best else None,
                             batched_gpix_per_s=best["gpix_per_s"] if best else None,
                             per_request_saturation_gpix_per_s=pr,
                             ratio=best["gpix_per_s"] / pr if best else None,
                             target="ratio >= 3 at p99 slowdown <= 2")
    return out


def config5_deep(ctx, M, WL, np, cpu=True):
    """Config 5: the whole 3840x2160-sample frame at p=1024 through the
    line-callback adapter (16 worker threads claiming rows, as MDZ's render
    threads do), then MDZ's MPFR loop on a seeded 64x64 crop on the host."""
    import threading
    v = WL.config5_view()
    W, H = v.real_width, v.real_height
    raw = np.full((H, W), -1, dtype=np.int32)
    fr = M.Frame(ctx, v, raw)
    nxt = [0]
    lock = threading.Lock()
    t0 = time.perf_counter()

    def worker():
        while True:
            with lock:
                line = nxt[0]
                nxt[0] += 1
                if line % 216 == 0:
                    log(f"config 5: row {line} at {time.perf_counter() - t0:.1f} s")
            if line >= H or fr.calculate_line(line) != 1:
                return

    ths = [threading.Thread(target=worker) for _ in range(16)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    el = time.perf_counter() - t0
    rc, st = fr.close()
    pi = float(np.where(raw > 0, raw, 20000).sum())
    out = dict(workload="config5 Mandelbrot 1920x1080 x 2x2 AA = 3840x2160 samples, p=1024, "
               "width 1e-290 at (-0.75, 0), depth 20000, via the line-callback adapter",
               rc=rc, seconds=el, pixel_iters=pi, gpix_per_s=pi / el / 1e9,
               sum_band_kernel_ms=st["kernel_ms"],
               frac_peak=2 * 2080 * pi / el / 1e12 / PEAK_TOPS, frac_peak_basis=FRAC_1024_BASIS,
               all_inside=bool((raw == 0).all()))
    if cpu:
        import random
        from oracle import oracle as O
        rng = random.Random(5)
        l0, c0 = rng.randrange(H - 64), rng.randrange(W - 64)
        rect = [s_.decode() for s_ in v._keep[:4]]
        try:
            O.use_reference(True)
            kind = "reference"
        except Exception:
            kind = "port"
        counts, cel = O.render(rect, precision=1024, depth=20000, width=W, height=H,
                               lines=(l0, 64, 1), cols=(c0, 64), threads=host_cores(),
                               timed=True)
        O.use_reference(False)
        cpi = float(np.where(counts > 0, counts, 20000).sum())
        out["cpu_baseline"] = dict(value=cpi / cel / 1e9, unit="G pixel-iter/s",
                                   cores=host_cores(), kind=kind,
                                   sample=f"64x64 crop at row {l0} col {c0}: {cpi:.3e} "
                                          f"pixel-iterations in {cel:.2f} s")
        out["crop_matches_gpu"] = bool(np.array_equal(counts, raw[l0:l0 + 64, c0:c0 + 64]))
    return out


def fp64_interior(ctx, WL, reps=5):
    """The FP64 kernel on a frame whose pixels all run to depth (1920x1080,
    depth 2000, 4.15e9 pixel-iterations): the iteration rate without refills.
    Its loop issues FP64_VALU_PER_ITER VALU instructions per wave-iteration,
    which bounds the rate at PEAK_VALU_ISSUE * 64 / FP64_VALU_PER_ITER."""
    v = WL.interior_view()
    ctx.render(v)
    ks = []
    for _ in range(reps):
        _, st = ctx.render(v)
        ks.append(st["kernel_ms"])
    k = sum(ks) / len(ks)
    rate = st["pixel_iters"] / (k * 1e-3)
    bound = PEAK_VALU_ISSUE * 64 / FP64_VALU_PER_ITER
    return dict(workload="1920x1080 inside the main cardioid (-0.1 + 0.1i, width 1e-6), depth 2000",
                pixel_iters=st["pixel_iters"], kernel_ms=k, kernel_gpix_per_s=rate / 1e9,
                frac_fp64_peak=7 * rate / 1e12 / PEAK_TOPS,
                valu_bound_gpix_per_s=bound / 1e9, frac_valu_bound=rate / bound)


FRAC_1024_BASIS = ("algorithmic 2080 32x32-bit MACs (x2 ops) per pixel-iteration: one 32-word "
                   "product and two squares; every pixel of this view lies within 1e-290 of the "
                   "real axis, where the loop bounds wim^2 by its exponent instead of squaring, "
                   "so the count includes a square that does not execute")


def rn1024_band(ctx, WL, rows):
    """Rows 1048.. of the config 5 frame (3840 samples each, p=1024, depth 20000)."""
    v5 = WL.config5_view()
    ctx.render(v5, lines=(1048, 2), cols=(0, 3840))
    _, st5 = ctx.render(v5, lines=(1048, rows), cols=(0, 3840))
    return dict(workload=f"config5 rows 1048-{1047 + rows} x 3840 samples, p=1024, depth 20000",
                pixel_iters=st5["pixel_iters"], kernel_ms=st5["kernel_ms"],
                kernel_gpix_per_s=st5["pixel_iters"] / (st5["kernel_ms"] * 1e-3) / 1e9,
                frac_peak=2 * 2080 * st5["pixel_iters"] / (st5["kernel_ms"] * 1e-3)
                / 1e12 / PEAK_TOPS, frac_peak_basis=FRAC_1024_BASIS)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the config 2 timed steps (for counter runs)")
    ap.add_argument("--band-only", type=int, default=0, metavar="ROWS",
                    help="only the 1024-bit figure over ROWS rows of config 5 (for counter runs)")
    ap.add_argument("--fp64-interior", action="store_true",
                    help="only the FP64 all-interior frame (for counter runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only a plain device-to-device copy of MIB MiB (calibrates "
                         "FETCH_SIZE / WRITE_SIZE against known bytes)")
    ap.add_argument("--op-probe", action="store_true",
                    help="only the per-operation probe of the RN arithmetic at 128 and 1024 "
                         "bits (for counter runs: one launch per operation, in OPP order)")
    ap.add_argument("--rate-probe", action="store_true",
                    help="only the issue rates of v_mad_u64_u32 and v_fma_f64 (the roofline "
                         "peak of the RN kernels)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the escape-time tiles shard across GPUs with no "
              "exchange, and this build runs on one device only (use --gpus 1).",
              file=sys.stderr)
        sys.exit(2)

    if a.rate_probe:
        import mdz_amd as M
        with M.Context(0) as ctx:
            out = {}
            for b in (1, 2, 4, 8):
                mad = ctx.rate_probe(0, b)
                fma = ctx.rate_probe(1, b)
                # lanes per CU per clock at the 2.4 GHz peak clock
                out[f"blocks_per_cu_{b}"] = dict(
                    mad_ops_per_s=mad, fma_ops_per_s=fma, mad_over_fma=mad / fma,
                    mad_lanes_per_cu_clk=mad / 256 / 2.4e9, fma_lanes_per_cu_clk=fma / 256 / 2.4e9)
                log(f"blocks/CU {b}: v_mad_u64_u32 {mad / 1e12:.2f} T lane-ops/s, "
                    f"v_fma_f64 {fma / 1e12:.2f} T, ratio {mad / fma:.3f}")
        print(json.dumps(out))
        return

    if a.op_probe:
        import mdz_amd as M
        from mdz_amd import workloads as WL
        names = ("loop_only", "product", "square", "add_sub", "squares_diff", "escape",
                 "iteration")
        out = {}
        with M.Context(0) as ctx:
            for leg, view, lines, warm, reps in (
                    ("rn128", WL.config2_view(), None, 1000, 4000),
                    ("rn1024", WL.config5_view(), (1048, 16), 200, 200)):
                H = lines[1] if lines else view.real_height
                W = view.real_width
                waves = (H * W + 63) // 64
                r = dict(workload=f"{W}x{H} pixels of the "
                                  f"{'config 2' if leg == 'rn128' else 'config 5'} view, "
                                  f"{warm} iterations first", waves=waves, warm=warm, reps=reps,
                         kernel_ms={})
                for op, nm in enumerate(names):
                    ms = ctx.op_probe(view, op, warm=warm, reps=reps, lines=lines)
                    r["kernel_ms"][nm] = ms
                    log(f"{leg} {nm}: {ms:.3f} ms")
                out[leg] = r
        print(json.dumps(out))
        return

    if a.copy_probe:
        import torch
        n = a.copy_probe << 18
        x = torch.ones(n, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps(dict(copy_probe_bytes=4 * n, launches=3)))
        return

    import numpy as np
    import mdz_amd as M
    from mdz_amd import workloads as WL

    ctx = M.Context(0)
    if a.band_only:
        r = rn1024_band(ctx, WL, a.band_only)
        ctx.close()
        print(json.dumps(r))
        return
    if a.fp64_interior:
        r = fp64_interior(ctx, WL)
        ctx.close()
        print(json.dumps(r))
        return
    res = dict(metric=METRIC, unit="G pixel-iter/s", n_gpus=1, higher_is_better=True,
               scaling="weak", vs_baseline=None, data="synthetic")

    if a.config == "2":
        v = WL.config2_view()
        log("config 2: warmup")
        for _ in range(a.warmup):
            ctx.render(v)
        kms, pis = [], []
        t0 = time.perf_counter()
        for i in range(a.steps):
            _, st = ctx.render(v)
            kms.append(st["kernel_ms"])
            pis.append(st["pixel_iters"])
        el = time.perf_counter() - t0
        log(f"config 2: {a.steps} steps in {el:.3f} s")
        pi_step = pis[0]
        kavg = sum(kms) / len(kms)
        ops = WL.ops_per_pixel_iter(M.PATH_MPFR, 128)
        achieved = 2.0 * ops * pi_step / (kavg * 1e-3) / 1e12
        pc = profiled_counters(CONFIG2_COUNTERS)
        roof = dict(bound="valu", achieved=achieved, peak=PEAK_TOPS, unit="TOP/s",
                    frac=achieved / PEAK_TOPS, traffic=pc["traffic_bytes"] if pc else None,
                    kernel_ms=kavg,
                    ops="2n^2+n = 36 algorithmic 32x32-bit MACs (x2 ops) per pixel-iteration "
                        "(n = 4) x 1.536e9 per launch")
        if pc:
            vi = pc["valu_insts_per_launch"]
            roof["counters"] = dict(source=pc["source"], build=pc.get("build"),
                                    note="rocprofv3 --pmc of the build named, read from the "
                                         "profile summary at run time")
            roof["valu_issue"] = dict(insts_per_launch=vi, achieved=vi / (kavg * 1e-3),
                                      peak=PEAK_VALU_ISSUE, unit="wave64 VALU instructions/s",
                                      frac=vi / (kavg * 1e-3) / PEAK_VALU_ISSUE)
        res.update(value=sum(pis) / el / 1e9, steps=a.steps, warmup=a.warmup,
                   ms_per_step=1e3 * el / a.steps, dtype="u32 limbs, RN p=128",
                   config=dict(workload="config2 Mandelbrot MPFR-path p=128 640x480 "
                               "centre (-0.743643887037158704752191506114774, "
                               "0.131825904205311970493132056385139) width 1e-25 depth 5000",
                               pixel_iters_per_step=pi_step, kernel="k_rn_persist<4,true>"),
                   kernel_gpix_per_s=pi_step / (kavg * 1e-3) / 1e9, roofline=roof)

        if a.headline_only:
            ctx.close()
            print(json.dumps(res))
            return

        # the same frame shape and arithmetic where pixels escape
        ve = WL.config2_escaping_view()
        ctx.render(ve)
        ke, we = [], []
        for _ in range(5):
            _, st = ctx.render(ve)
            ke.append(st["kernel_ms"])
            we.append(st["wall_ms"])
        kea = sum(ke) / len(ke)
        res["rn128_escaping"] = dict(
            workload="640x480 p=128 depth 5000, 1e-24 wide at 0.2926 + 0.4557i (37 % of pixels "
                     "escape)", pixel_iters=st["pixel_iters"], kernel_ms=kea,
            kernel_gpix_per_s=st["pixel_iters"] / (kea * 1e-3) / 1e9,
            wall_gpix_per_s=st["pixel_iters"] / (float(np.median(we)) * 1e-3) / 1e9)
        log(f"rn128 escaping view: {res['rn128_escaping']['kernel_gpix_per_s']:.1f} G "
            f"pixel-iter/s (kernel)")

        # FP64 figure: config 1 frame
        v1 = WL.config1_view()
        for _ in range(3):
            ctx.render(v1)
        k1, w1, p1 = [], [], 0
        for _ in range(50):
            _, st = ctx.render(v1)
            k1.append(st["kernel_ms"])
            w1.append(st["wall_ms"])
            p1 = st["pixel_iters"]
        k1a = sum(k1) / len(k1)
        res["fp64"] = dict(workload="config1 240x180 depth 1000", pixel_iters=p1,
                           kernel_ms=k1a, wall_ms_p50=float(np.median(w1)),
                           kernel_gpix_per_s=p1 / (k1a * 1e-3) / 1e9,
                           frac_fp64_peak=7 * p1 / (k1a * 1e-3) / 1e12 / PEAK_TOPS)
        log(f"fp64: {res['fp64']['kernel_gpix_per_s']:.1f} G pixel-iter/s (kernel)")
        # the long double path (RN at 64 bits) on the same frame
        vld = WL.config1_view(path=M.PATH_LD)
        for _ in range(3):
            ctx.render(vld)
        kl = []
        for _ in range(20):
            _, st = ctx.render(vld)
            kl.append(st["kernel_ms"])
        kla = sum(kl) / len(kl)
        res["fp64"]["ld_path"] = dict(workload="config1 240x180 depth 1000, long double path "
                                      "(round-to-nearest at 64 bits)",
                                      pixel_iters=st["pixel_iters"], kernel_ms=kla,
                                      kernel_gpix_per_s=st["pixel_iters"] / (kla * 1e-3) / 1e9)
        # the FP64 kernel on a frame that fills the device: the same view at 3840x2160
        v4 = WL.config1_view(W=3840, H=2160)
        ctx.render(v4)
        k4 = []
        for _ in range(5):
            _, st = ctx.render(v4)
            k4.append(st["kernel_ms"])
        k4a = sum(k4) / len(k4)
        p4 = st["pixel_iters"]
        res["fp64"]["at_3840x2160"] = dict(pixel_iters=p4, kernel_ms=k4a,
                                           kernel_gpix_per_s=p4 / (k4a * 1e-3) / 1e9,
                                           frac_fp64_peak=7 * p4 / (k4a * 1e-3) / 1e12 / PEAK_TOPS)
        log(f"fp64 3840x2160: {res['fp64']['at_3840x2160']['kernel_gpix_per_s']:.1f} G "
            f"pixel-iter/s (kernel)")
        res["fp64"]["interior_1920x1080"] = fp64_interior(ctx, WL)
        log(f"fp64 interior: {res['fp64']['interior_1920x1080']['kernel_gpix_per_s']:.1f} G "
            f"pixel-iter/s (kernel)")

        # 1024-bit figure: a band of config 5 rows
        res["rn1024"] = rn1024_band(ctx, WL, 192)
        log(f"rn1024: {res['rn1024']['kernel_gpix_per_s']:.3f} G pixel-iter/s (kernel)")

        # p50/p99 tile latency: continuous batching, 64x64 config 2 crops,
        # against one launch per request (the serving target: >= 3x the
        # per-request form's throughput at p99 <= 2x unloaded)
        tiles, _ = crop_tiles(WL, 300)
        unl = unloaded_latency(ctx, M, np, M.SERVE_BATCHED, tiles)
        sat = saturation_run(ctx, M, np, M.SERVE_BATCHED, tiles)
        unl_pr = unloaded_latency(ctx, M, np, M.SERVE_PER_REQUEST, tiles)
        sat_pr = saturation_run(ctx, M, np, M.SERVE_PER_REQUEST, tiles)
        loads = []
        for f in (0.5, 0.6, 0.7, 0.8, 0.9):
            r = serve_run(ctx, M, np, M.SERVE_BATCHED, tiles, f * sat["tiles_per_s"])
            loads.append(dict(load=f, rate=r["rate"], p50_ms=r["p50_ms"], p99_ms=r["p99_ms"],
                              tiles_per_s=r["tiles_per_s"], gpix_per_s=r["gpix_per_s"],
                              p99_over_unloaded=r["p99_ms"] / unl))
        ok = [x for x in loads if x["p99_over_unloaded"] <= 2.0]
        best = max(ok, key=lambda x: x["gpix_per_s"]) if ok else None
        half = loads[0]
        res["latency"] = dict(workload="Poisson requests of 64x64 crops of config 2 (p=128, "
                              "depth 5000), continuous batching", unloaded_p50_ms=unl,
                              saturation_tiles_per_s=sat["tiles_per_s"],
                              saturation_gpix_per_s=sat["gpix_per_s"],
                              at_half_load=dict(rate=half["rate"], p50_ms=half["p50_ms"],
                                                p99_ms=half["p99_ms"],
                                                tiles_per_s=half["tiles_per_s"]),
                              sweep=loads,
                              per_request=dict(unloaded_p50_ms=unl_pr,
                                               saturation_tiles_per_s=sat_pr["tiles_per_s"],
                                               saturation_gpix_per_s=sat_pr["gpix_per_s"]),
                              batching_target=dict(
                                  load=best["load"] if best else None,
                                  p99_over_unloaded=best["p99_over_unloaded"] if best else None,
                                  batched_gpix_per_s=best["gpix_per_s"] if best else None,
                                  per_request_saturation_gpix_per_s=sat_pr["gpix_per_s"],
                                  ratio=best["gpix_per_s"] / sat_pr["gpix_per_s"] if best else None,
                                  target="ratio >= 3 at p99 <= 2x unloaded latency"))
        log(f"latency: unloaded {unl:.2f} ms; half load p50 {half['p50_ms']:.2f} "
            f"p99 {half['p99_ms']:.2f} ms; per-request saturation {sat_pr['gpix_per_s']:.1f} G; "
            f"batching target ratio {res['latency']['batching_target']['ratio']}")

        # config 3: 256 mixed tiles, persistent batch against one launch per tile
        c3 = config3_batched(ctx, M, WL, np)
        res["config3"] = dict(workload=c3["config"]["workload"], **c3["batched"])

        if not a.no_cpu_baseline:
            log("cpu baseline (MDZ's loops on host cores) ...")
            th = host_cores()
            rect = [s.decode() for s in v._keep[:4]]
            legs = {}
            # 128 bits: config 2, every 4th row
            for b in ("O0", "O2"):
                legs[f"rn128_{b}"] = cpu_baseline(rect, 128, 5000, 640, 480, 4, th, build=b)
                log(f"cpu 128-bit {b}: {legs[f'rn128_{b}']['value']:.4f} G pixel-iter/s")
            # 1024 bits: a seeded 64x64 crop of config 5
            import random
            v5 = WL.config5_view()
            rng = random.Random(5)
            l0, c0 = rng.randrange(v5.real_height - 64), rng.randrange(v5.real_width - 64)
            r5 = [s_.decode() for s_ in v5._keep[:4]]
            for b in ("O0", "O2"):
                legs[f"rn1024_{b}"] = cpu_baseline(r5, 1024, 20000, v5.real_width,
                                                   v5.real_height, 1, th, build=b,
                                                   lines=(l0, 64, 1), cols=(c0, 64),
                                                   sample_name="config 5")
                log(f"cpu 1024-bit {b}: {legs[f'rn1024_{b}']['value']:.4f} G pixel-iter/s")
            # long double (MDZ's "double" path) and the IEEE double loop: config 1
            r1 = [s_.decode() for s_ in v1._keep[:4]]
            for b in ("O0", "O2"):
                legs[f"ld_{b}"] = cpu_baseline(r1, 64, 1000, 240, 180, 1, th, path=M.PATH_LD,
                                               build=b, sample_name="config 1", repeat=30)
                log(f"cpu long double {b}: {legs[f'ld_{b}']['value']:.3f} G pixel-iter/s")
            legs["double_O2"] = cpu_baseline(r1, 64, 1000, 240, 180, 1, th,
                                             path=M.PATH_DOUBLE, sample_name="config 1",
                                             repeat=30)
            log(f"cpu double: {legs['double_O2']['value']:.3f} G pixel-iter/s")
            gpu = dict(rn128=res["value"], rn1024=res["rn1024"]["kernel_gpix_per_s"],
                       ld=res["fp64"]["ld_path"]["kernel_gpix_per_s"],
                       double=res["fp64"]["kernel_gpix_per_s"])
            ratios = {k: gpu[k.rsplit("_", 1)[0]] / x["value"] for k, x in legs.items()}
            res["cpu_baseline"] = dict(legs["rn128_O0"])
            res["cpu_baseline"]["legs"] = legs
            res["cpu_baseline"]["gpu_over_cpu"] = ratios
            res["cpu_baseline"]["gpu_figures"] = gpu
            log(f"cpu baseline: {res['cpu_baseline']['value']:.4f} G pixel-iter/s")
    elif a.config == "3":
        res.update(config3_batched(ctx, M, WL, np, classes=True, graphs_off=True))
    elif a.config == "gmp":
        res.update(gmp_path(ctx, M, WL, a, cpu=not a.no_cpu_baseline))
    elif a.config == "colour":
        r = colour_pass(ctx, M, np, a, cpu=not a.no_cpu_baseline)
        res.update(metric="M output pixels/s of the colouring pass (3840x2160, aa=2)",
                   unit="M output pixels/s", value=r["aa2"]["mpix_per_s"], steps=a.steps,
                   warmup=2, ms_per_step=r["aa2"]["kernel_ms"], dtype="f64 colour arithmetic",
                   config=dict(workload="colouring 3840x2160, 2x2 anti-aliasing, 256-colour "
                               "interpolated palette", kernel="k_colour"),
                   roofline=r["aa2"]["roofline"], colour=r)
        if "cpu_baseline" in r:
            res["cpu_baseline"] = r["cpu_baseline"]
    elif a.config == "5":
        r = config5_deep(ctx, M, WL, np, cpu=not a.no_cpu_baseline)
        res.update(value=r["gpix_per_s"], steps=1, warmup=0, ms_per_step=1e3 * r["seconds"],
                   dtype="u32 limbs, RN p=1024", config=dict(workload=r["workload"]), deep=r)
        if "cpu_baseline" in r:
            res["cpu_baseline"] = r["cpu_baseline"]
    elif a.config == "4":
        sv = config4_serving(ctx, M, WL, np)
        b, t = sv["batched"], sv["per_request"]
        res.update(value=b["saturation"]["gpix_per_s"], steps=1, warmup=0,
                   ms_per_step=None, dtype="u32 limbs, RN p=128",
                   config=dict(workload="config4 Poisson requests of 64x64 crops of the config 2 "
                               "view (p=128, depth 5000), swept to saturation"),
                   serving=sv,
                   throughput_ratio_batched_vs_per_request=b["saturation"]["gpix_per_s"] /
                   t["saturation"]["gpix_per_s"])
        res["serving_config3_mix"] = config4_mix(ctx, M, WL, np)
    else:
        raise SystemExit(f"unknown --config {a.config}")

    ctx.close()
    print(json.dumps(res))


if __name__ == "__main__":
    main()
