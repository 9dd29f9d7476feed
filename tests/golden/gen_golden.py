"""Write the golden count maps (tests/golden/golden.npz + golden.json).

Counts come from MDZ's own formula loops (src/frac_*.c compiled as they are
into oracle/_ref/libfrac_ref.so by oracle/Makefile), driven per pixel by the
oracle's restatement of the line workers (src/fractal.c:29-257).  Run in a
container that has the reference tree:

    make -C oracle all && python tests/golden/gen_golden.py

Deep views.  Besides the configurations and gallery scenes, the maps cover
every formula at every precision class of the device kernels (1 to 32 words,
word-aligned and ragged p), the long double and double paths and the GMP path,
at widths down to 1e-290, each a crop whose pixels escape at many different
iterations.  A view is found by bisection on a segment from a point that
does not escape within the depth (the origin, or the Misiurewicz point c = i
of the Mandelbrot set, whose orbit i, -1+i, -i, -1+i, ... is exact) to a point
that escapes: the bisection stops once the segment is 1/64 of the width, and
the view is centred on its midpoint, so it straddles the boundary of the set
of pixels that escape within the depth.  Orbits of neighbouring pixels there
separate exponentially, so the counts spread over hundreds to thousands of
iterations.  The search runs the oracle's restated loops (pinned equal to
MDZ's compiled loops); the maps themselves come from MDZ's compiled loops.
Every crop except the three throughput configurations, whose views lie
inside the set by construction, must hold at least 20 distinct counts.
"""
import concurrent.futures as cf
import json
import math
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from tests import cases  # noqa: E402
from mdz_amd import workloads as WL  # noqa: E402  (pure-Python view descriptions)

OUT = os.path.dirname(os.path.abspath(__file__))


def vdict(rect, precision, depth, W, H, fractal=0, family=0, path=0, julia=(None, None)):
    return dict(rect=list(rect), precision=precision, depth=depth, width=W, height=H,
                fractal=fractal, family=family, path=path, julia=list(julia))


def cases_list():
    out = []
    # config 1 (double and long double paths), every 4th row
    v1 = WL.config1_view()
    r1 = [s.decode() for s in v1._keep[:4]]
    for path in (O.PATH_DOUBLE, O.PATH_LD):
        for fr in range(4):
            out.append((f"config1_{'double' if path == 2 else 'ld'}_f{fr}",
                        vdict(r1, 64, 1000, 240, 180, fractal=fr, path=path),
                        (0, 45, 4), (0, 240)))
    # gallery scenes
    for name in cases.SCENES:
        rect, p, depth, fr, path, (W, H) = cases.scene_rect(name)
        ch, cw = min(32, H), 48
        out.append((f"scene_{name}", vdict(rect, p, depth, W, H, fractal=fr, path=path),
                    (H // 2 - ch // 2, ch, 1), (W // 2 - cw // 2, cw)))
    # Julia family
    J = cases.JULIA
    for fr in range(4):
        # the Burning Ship's Julia set at c = -0.8 + 0.156i is dust (every
        # pixel escapes within 10 iterations): its map uses c = -1.75 - 0.03i
        jc = J["c"] if fr != 1 else ("-1.75", "-0.03")
        out.append((f"julia_f{fr}", vdict(J["rect"], J["precision"], J["depth"], 96, 72,
                                          fractal=fr, family=1, julia=jc),
                    (0, 24, 3), (0, 96)))
    # config 2 and config 5 crops
    v2 = WL.config2_view()
    out.append(("config2_crop", vdict([s.decode() for s in v2._keep[:4]], 128, 5000, 640, 480),
                (224, 8, 1), (300, 16)))
    v5 = WL.config5_view()
    out.append(("config5_crop", vdict([s.decode() for s in v5._keep[:4]], 1024, 20000, 3840, 2160),
                (1076, 4, 1), (1916, 8)))
    # config 3: crops of the first tiles with depth <= 4000, and of the first
    # two tiles of the FP64 and 64-bit classes
    n = 0
    seen = {53: 0, 64: 0}
    for sp in WL.config3_specs():
        small = sp["depth"] <= 4000 and sp["precision"] != 53 and n < 4
        cls = sp["precision"] in seen and seen[sp["precision"]] < 2
        if not (small or cls):
            continue
        if small:
            n += 1
        else:
            seen[sp["precision"]] += 1
        v = WL.spec_view(sp)
        out.append((f"config3_tile{sp['index']}",
                    vdict([s.decode() for s in v._keep[:4]], v.precision, sp["depth"],
                          sp["side"], sp["side"], fractal=sp["fractal"], path=v.path),
                    (sp["side"] // 2 - 4, 8, 1), (sp["side"] // 2 - 4, 8)))
    return out


# ---- deep views -----------------------------------------------------------
# (precision, path, depth, log10 width) per precision class: the word count N
# of the device kernel is ceil(p / 32) rounded up to 1, 2, 3, 4, 6, 8, 12, 16,
# 24 or 32; p = 176, 416 and 1000 are ragged (rounded inside a word)
DEEP_CLASSES = [
    (32, O.PATH_MPFR, 2000, -3), (64, O.PATH_MPFR, 20000, -14), (96, O.PATH_MPFR, 5000, -17),
    (128, O.PATH_MPFR, 5000, -25), (176, O.PATH_MPFR, 10000, -40),
    (192, O.PATH_MPFR, 10000, -45), (256, O.PATH_MPFR, 10000, -60),
    (384, O.PATH_MPFR, 20000, -95), (416, O.PATH_MPFR, 20000, -100),
    (512, O.PATH_MPFR, 20000, -100), (768, O.PATH_MPFR, 20000, -200),
    (1000, O.PATH_MPFR, 20000, -250), (1024, O.PATH_MPFR, 20000, -290),
    (64, O.PATH_LD, 20000, -14), (64, O.PATH_DOUBLE, 20000, -11),
    (128, O.PATH_GMP, 5000, -25), (1024, O.PATH_GMP, 20000, -250),
]
DEEP_SIDE = 32


def deep_segment(fractal, precision):
    """(a, b, side): a does not escape, b escapes quickly.  Rays from the
    origin at 1.2 rad (Mandelbrot, Burning Ship) and 2.6 rad (Celtic, Variant);
    from 768 bits (widths from 1e-200 down) the Mandelbrot ray meets the set
    where neighbouring orbits separate too slowly to escape apart within the
    depth, so the segment runs from c = i instead, with a 64-pixel crop."""
    if fractal == 0 and precision >= 768:
        return 1j, 0.05 + 1j, 64
    th = 1.0 if (fractal == 0 and precision == 128) else (1.2 if fractal <= 1 else 2.6)
    if precision >= 1000 and fractal >= 2:
        th = 2.5 if fractal == 2 else 2.7
    return 0j, 2.5 * complex(math.cos(th), math.sin(th)), DEEP_SIDE


def _hexs(X, K):
    return f"{'-' if X < 0 else ''}0x{abs(X):x}p-{K}"


def _escapes(x, y, K, precision, depth, fractal, path):
    """c = (x, y) * 2^-K exactly: the pixel (0, 0) of a 1 x 1 view whose
    corner is c (src/fractal.c:159-188 leaves x = xmin, y = ymax there)."""
    w = 1 << (K - 4)
    rect = (_hexs(x, K), _hexs(x + w, K), _hexs(y, K), _hexs(w, K))
    return int(O.render(rect, precision=precision, depth=depth, width=1, height=1,
                        fractal=fractal, path=path)[0, 0]) != 0


def find_deep_centre(fractal, precision, path, depth, log10w):
    """Bisection from a (inside within depth) to b (escapes): the centre as
    exact hex strings."""
    a, b, side = deep_segment(fractal, precision)
    K = max(precision, 80) + 16
    fx = lambda v: int(round(v * 2**50)) << (K - 50)
    ax, ay, bx, by = fx(a.real), fx(a.imag), fx(b.real), fx(b.imag)
    assert not _escapes(ax, ay, K, precision, depth, fractal, path)
    assert _escapes(bx, by, K, precision, depth, fractal, path)
    stop = 1 << (K + int(math.floor(log10w * math.log2(10))) - 6)
    while max(abs(bx - ax), abs(by - ay)) > stop:
        mx, my = (ax + bx) >> 1, (ay + by) >> 1
        if _escapes(mx, my, K, precision, depth, fractal, path):
            bx, by = mx, my
        else:
            ax, ay = mx, my
    return _hexs((ax + bx) >> 1, K), _hexs((ay + by) >> 1, K), side


def _deep_job(args):
    fractal, precision, path, depth, log10w = args
    O.use_reference(False)        # the restated loops, pinned equal to MDZ's
    cx, cy, side = find_deep_centre(fractal, precision, path, depth, log10w)
    rect = O.rect_from_center(cx, cy, f"1e{log10w}", precision, side, side,
                              coords=(path == O.PATH_GMP))
    tag = {O.PATH_MPFR: "mpfr", O.PATH_LD: "ld", O.PATH_DOUBLE: "double", O.PATH_GMP: "gmp"}[path]
    name = f"deep_{tag}{precision}_f{fractal}"
    return (name, vdict(rect, precision, depth, side, side, fractal=fractal, path=path),
            (0, side, 1), (0, side))


def deep_cases():
    jobs = [(fr, p, path, depth, lw) for (p, path, depth, lw) in DEEP_CLASSES
            for fr in range(4)]
    with cf.ProcessPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        return list(ex.map(_deep_job, jobs))


# views inside the set by construction (configs 2, 3, 5: throughput cases)
ALL_INSIDE_OK = ("config2_crop", "config5_crop", "config3_tile")


def main():
    O.build()
    O.lib()
    todo = cases_list() + deep_cases()
    O.use_reference(True)          # MDZ's own loops
    arrays, meta = {}, []
    for name, vd, lines, cols in todo:
        counts = O.render(vd["rect"], precision=vd["precision"], depth=vd["depth"],
                          width=vd["width"], height=vd["height"], fractal=vd["fractal"],
                          family=vd["family"], path=vd["path"],
                          julia=tuple(vd["julia"]), lines=lines, cols=cols,
                          threads=os.cpu_count() or 4)
        arrays[name] = counts.astype(np.int32)
        meta.append(dict(name=name, view=vd, lines=list(lines), cols=list(cols)))
        nd = len(np.unique(counts))
        print(f"{name}: {counts.shape} distinct={nd} min={counts[counts > 0].min(initial=np.iinfo(counts.dtype).max)} "
              f"max={counts.max()} zeros={(counts == 0).mean():.2f}", flush=True)
        assert nd >= 20 or name.startswith(ALL_INSIDE_OK), f"{name}: {nd} distinct counts"
    np.savez_compressed(os.path.join(OUT, "golden.npz"), **arrays)
    with open(os.path.join(OUT, "golden.json"), "w") as f:
        json.dump(dict(source="src/frac_*.c compiled as-is (oracle/_ref), -O0 -ggdb -std=gnu99",
                       cases=meta), f, indent=1)


if __name__ == "__main__":
    main()
