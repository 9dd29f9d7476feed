"""Synthetic workloads of the benchmark configurations (seeded, reproducible).

  config 1  FP64 Mandelbrot 240x180, centre (-0.75, 0), width 3.0, depth 1000,
            pixel-centre convention c = centre + ((x - W/2) + 1/2, (H/2 - y) - 1/2) * width/W,
            expressed as MDZ's corner rect (xmin = cx - (W-1)/2 * s, ymax = cy + (H-1)/2 * s)
  config 2  MPFR p=128 Mandelbrot 640x480 at the seahorse-valley centre, width 1e-25, depth 5000
  config 3  256 seeded tiles: side in {64,128,256,512}, log10(width) ~ U[-250,-2],
            centre (-0.75, 0) + offsets of up to 1e-3*width, precision class
            = smallest of {FP64,64,128,256,512,1024} >= ceil(log2(4*side/width)) + 32,
            depth log-uniform in [1e3, 5e4], formulas 60/20/20 Mandelbrot/Burning Ship/Celtic
  config 4  Poisson arrivals of config-3 tiles (the serving loop)
  config 5  Mandelbrot 1920x1080 with 2x2 anti-aliasing (3840x2160 samples), p=1024,
            width 1e-290 at (-0.75, 0), depth 20000

Views of configs 2, 3 and 5 are built the way MDZ's command-line path builds
them from a centre and size (mdz_amd.rect_from_center).
"""
import math
import random
from decimal import Decimal, getcontext

from . import (BURNING_SHIP, GENERALIZED_CELTIC, MANDELBROT, PATH_DOUBLE, PATH_MPFR,
               make_view, rect_from_center)

CONFIG2_CENTRE = ("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139")
PREC_CLASSES = (53, 64, 128, 256, 512, 1024)     # 53 = the FP64 class
CONFIG3_SEED = 0x4D445A


def config1_view(fractal=MANDELBROT, path=PATH_DOUBLE, W=240, H=180):
    width = Decimal(3)
    s = width / W
    xmin = Decimal("-0.75") - Decimal(W - 1) / 2 * s
    ymax = Decimal(0) + Decimal(H - 1) / 2 * s
    rect = (str(xmin), str(xmin + width), str(ymax), str(width))
    return make_view(rect, precision=64, depth=1000, width=W, height=H, fractal=fractal,
                     path=path)


def interior_view(path=PATH_DOUBLE):
    """A 1920x1080 frame inside the main cardioid (centre -0.1 + 0.1i, width
    1e-6, depth 2000): every pixel runs to depth, so the frame measures the
    iteration rate alone."""
    rect = rect_from_center("-0.1", "0.1", "1e-6", 64, 1920, 1080)
    return make_view(rect, precision=64, depth=2000, width=1920, height=1080, path=path)


def config2_view():
    rect = rect_from_center(CONFIG2_CENTRE[0], CONFIG2_CENTRE[1], "1e-25", 128, 640, 480)
    return make_view(rect, precision=128, depth=5000, width=640, height=480)


# a point on the boundary near 0.2926 + 0.4557i (found by bisection between a
# point inside and one that escapes): a 640x480 frame 1e-24 wide there at
# p=128 and depth 5000 has 37 % of its pixels escaping, at hundreds of
# distinct counts, most of them late in the orbit
ESCAPING_CENTRE = ("0.292609052674571593299537901194696179",
                   "0.455711598939330790877439942797882725")


def config2_escaping_view():
    """The config 2 frame shape and arithmetic (640x480, p=128, depth 5000) at
    a boundary point where pixels escape, so refills and lanes that finish
    early are part of the measurement."""
    rect = rect_from_center(ESCAPING_CENTRE[0], ESCAPING_CENTRE[1], "1e-24", 128, 640, 480)
    return make_view(rect, precision=128, depth=5000, width=640, height=480)


def config5_view(aa=2):
    W, H = 1920, 1080
    rect = rect_from_center("-0.75", "0", "1e-290", 1024, W, H)
    return make_view(rect, precision=1024, depth=20000, width=W * aa, height=H * aa)


def precision_class(side, log10_width):
    need = math.ceil(math.log2(4 * side) - log10_width * math.log2(10)) + 32
    for p in PREC_CLASSES:
        if p >= need:
            return p
    return PREC_CLASSES[-1]


def config3_specs(n=256, seed=CONFIG3_SEED):
    """Tile descriptions (dicts) of config 3, deterministic in `seed`."""
    rng = random.Random(seed)
    getcontext().prec = 400
    out = []
    for i in range(n):
        side = rng.choice((64, 128, 256, 512))
        lw = rng.uniform(-250.0, -2.0)
        e = math.floor(lw)
        mant = 10 ** (lw - e)
        width = Decimal(f"{mant:.15f}") * Decimal(10) ** e
        dx, dy = rng.uniform(-1, 1), rng.uniform(-1, 1)
        cx = Decimal("-0.75") + Decimal(f"{dx:.15f}") * Decimal("1e-3") * width
        cy = Decimal(0) + Decimal(f"{dy:.15f}") * Decimal("1e-3") * width
        depth = int(round(10 ** rng.uniform(3.0, math.log10(5e4))))
        u = rng.random()
        fractal = MANDELBROT if u < 0.6 else (BURNING_SHIP if u < 0.8 else GENERALIZED_CELTIC)
        p = precision_class(side, lw)
        out.append(dict(index=i, side=side, log10_width=lw, cx=format(cx, "e"),
                        cy=format(cy, "e"), width=format(width, "e"), depth=depth,
                        fractal=fractal, precision=p))
    return out


def spec_view(sp):
    """View of one config-3 tile description."""
    p = sp["precision"]
    side = sp["side"]
    fp64 = p == 53
    rect = rect_from_center(sp["cx"], sp["cy"], sp["width"], 64 if fp64 else p, side, side)
    return make_view(rect, precision=64 if fp64 else p, depth=sp["depth"], width=side,
                     height=side, fractal=sp["fractal"],
                     path=PATH_DOUBLE if fp64 else PATH_MPFR)


def ops_per_pixel_iter(path, precision):
    """Algorithmic operations per pixel-iteration: 7 FP64 ops on the double path;
    2n^2+n 32x32-bit multiply-accumulates with n = ceil(p/32) on the RN paths
    (one general product n^2 and two squares n(n+1)/2)."""
    if path == PATH_DOUBLE:
        return 7
    n = -(-(64 if path == 1 else precision) // 32)
    return 2 * n * n + n
