"""Autotune every kernel class on this GPU and write the table mdzg_create loads.

    python -m mdz_amd.autotune [--out mdz_amd/tune_gfx950.txt] [--reps 2]

For each kernel id (FP64; the RN kernels by word count, word-aligned and
ragged precision; the GMP storage buckets) mdzg_autotune renders probe
frames at every pair of resident blocks per CU and refill chunk and keeps
the pair of least mean relative time.  A class has two probes, the two kinds
of deep render:

  * a view at the Misiurewicz point c = i of the Mandelbrot set (orbit i,
    -1+i, -i, -1+i, ... exactly), at a width matched to the precision: every
    pixel escapes, after a number of iterations that grows with the depth of
    the zoom and varies across the frame, so refills and uneven lanes cost
    what they cost in deep renders with structure;
  * a 640 x 480 view inside the set at the parabolic point -0.75 (every pixel
    runs to depth 2000, as in configs 2 and 5), where every lane holds one
    pixel from start to end.

The FP64 probes are the whole set (centre -0.75, width 3) at 240 x 180
(config 1) and 3840 x 2160, and a frame inside the main cardioid.  Then the
line adapter's band size and bands in flight are chosen on a 1024-bit frame
rendered through it by 16 worker threads, as MDZ's render threads call it.
"""
import argparse
import json
import os
import sys
import threading
import time

import numpy as np

import mdz_amd as M
from mdz_amd import workloads as WL

HERE = os.path.dirname(os.path.abspath(__file__))
NWORDS = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)
PROBE_W, PROBE_H = 1024, 512


def misiurewicz_view(precision, path=M.PATH_MPFR, W=PROBE_W, H=PROBE_H, depth=50000):
    """A view centred on c = i whose pixel spacing is about 2^-(p - 12)."""
    lw = -(precision - 12) * 0.30103 + np.log10(W)
    lw = min(lw, -1.0)
    e = int(np.floor(lw))
    size = f"{10 ** (lw - e):.6f}e{e}"
    rect = M.rect_from_center("0", "1", size, precision, W, H, coords=(path == M.PATH_GMP))
    return M.make_view(rect, precision=precision, depth=depth, width=W, height=H, path=path)


def inside_view(precision, path=M.PATH_MPFR, W=640, H=480, depth=2000):
    """A view at -0.75 + 0i, 1e-7 wide: escape times there grow like pi over
    the distance from -0.75, so every pixel runs to depth."""
    rect = M.rect_from_center("-0.75", "0", "1e-7", precision, W, H,
                              coords=(path == M.PATH_GMP))
    return M.make_view(rect, precision=precision, depth=depth, width=W, height=H, path=path)


def probes():
    """(kernel id, description, [views]) for every kernel class."""
    out = [(0, "FP64 240x180 + 3840x2160 + interior",
            [WL.config1_view(), WL.config1_view(W=3840, H=2160), WL.interior_view()])]
    for n in NWORDS:
        for full in (1, 0):
            p = 32 * n if full else max(2, 32 * n - 16)
            kid = M.kernel_id(M.PATH_MPFR, p)
            out.append((kid, f"RN {n} words p={p}",
                        [misiurewicz_view(max(p, 8)), inside_view(p)]))
    seen = set()
    for p in range(1024, 7, -8):         # the deepest precision of each GMP bucket
        kid = M.kernel_id(M.PATH_GMP, p)
        if kid in seen:
            continue
        seen.add(kid)
        out.append((kid, f"GMP p={p}", [misiurewicz_view(p, path=M.PATH_GMP),
                                        inside_view(p, path=M.PATH_GMP)]))
    return out


def frame_time(ctx, view, threads=16):
    raw = np.full((view.real_height, view.real_width), -1, dtype=np.int32)
    fr = M.Frame(ctx, view, raw)
    nxt = [0]
    lock = threading.Lock()

    def worker():
        while True:
            with lock:
                line = nxt[0]
                nxt[0] += 1
            if line >= view.real_height or fr.calculate_line(line) != 1:
                return

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(threads)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    el = time.perf_counter() - t0
    fr.close()
    return el


def tune_bands(ctx, reps=2):
    v = misiurewicz_view(1024, W=3840, H=128, depth=20000)
    frame_time(ctx, v)
    best = None
    res = []
    for bp in (1 << 16, 1 << 17, 1 << 18, 1 << 19):
        for inf in (2, 3, 4):
            ctx.option_set(M.OPT_BAND_PIXELS, bp)
            ctx.option_set(M.OPT_BAND_INFLIGHT, inf)
            t = min(frame_time(ctx, v) for _ in range(reps))
            res.append(dict(band_pixels=bp, inflight=inf, seconds=t))
            if best is None or t < best[0]:
                best = (t, bp, inf)
    ctx.option_set(M.OPT_BAND_PIXELS, best[1])
    ctx.option_set(M.OPT_BAND_INFLIGHT, best[2])
    return dict(best=dict(band_pixels=best[1], inflight=best[2], seconds=best[0]), sweep=res)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "tune_gfx950.txt"))
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--only", type=int, default=None, help="tune one kernel id")
    ap.add_argument("--no-bands", action="store_true")
    a = ap.parse_args()
    report = []
    env = os.environ.pop("MDZG_TUNE_TABLE", None)
    with M.Context(0) as ctx:
        for kid, desc, views in probes():
            if a.only is not None and kid != a.only:
                continue
            t0 = time.perf_counter()
            r = ctx.autotune(views, reps=a.reps)
            r["probe"] = desc
            r["seconds"] = time.perf_counter() - t0
            report.append(r)
            print(f"kernel {kid:2d} {desc:24s} blocks/CU {r['blocks_per_cu']} chunk "
                  f"{r['chunk']:4d}: probes {r['best_ms']:.3f} ms (was {r['default_ms']:.3f}, "
                  f"{r['candidates']} pairs, {r['seconds']:.1f} s)", file=sys.stderr, flush=True)
        bands = None if a.no_bands else tune_bands(ctx)
        if bands:
            print(f"bands: {bands['best']}", file=sys.stderr, flush=True)
        ctx.tune_save(a.out)
    if env is not None:
        os.environ["MDZG_TUNE_TABLE"] = env
    print(json.dumps(dict(kernels=report, bands=bands)))


if __name__ == "__main__":
    main()
