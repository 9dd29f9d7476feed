"""GPU parity: the HIP kernels, called through the C ABI, against the oracle
and the golden maps made from MDZ's own loops.  Tolerance: none (uint32
count maps must be bit-identical)."""
import random
import threading

import numpy as np
import pytest

import mdz_amd as M
from mdz_amd import workloads as WL
from tests import cases

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return np.array_equal(np.asarray(a).astype(np.int64), np.asarray(b).astype(np.int64))


def _rect(v):
    return [s.decode() for s in v._keep[:4]]


@pytest.mark.parametrize("path", [M.PATH_DOUBLE, M.PATH_LD])
@pytest.mark.parametrize("fractal", [0, 1, 2, 3])
def test_config1_full_frame(ctx, oracle, path, fractal):
    v = WL.config1_view(fractal=fractal, path=path)
    got, st = ctx.render(v)
    ref = oracle.render(_rect(v), precision=64, depth=1000, width=240, height=180,
                        fractal=fractal, path=path, threads=16)
    assert _eq(got, ref)
    assert st["pixel_iters"] == float(np.where(ref > 0, ref, 1000).sum())


def _golden_view(vd):
    return M.make_view(vd["rect"], precision=vd["precision"], depth=vd["depth"],
                       width=vd["width"], height=vd["height"], fractal=vd["fractal"],
                       family=vd["family"], path=vd["path"], julia=tuple(vd["julia"]))


def _golden_meta():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                           "golden.json")) as f:
        return json.load(f)["cases"]


DEEP = [c["name"] for c in _golden_meta() if c["name"].startswith("deep_")]


@pytest.mark.parametrize("name", DEEP)
def test_deep_golden_map(ctx, golden, name):
    """Boundary crops at every precision class, formula and path, down to
    width 1e-290 (p = 1024, depth 20000), whose pixels escape at many
    different iterations: MDZ's own loops reproduced on the device."""
    meta, expect = next((m, e) for m, e in golden if m["name"] == name)
    v = _golden_view(meta["view"])
    print(name, flush=True)
    got, st = ctx.render(v)
    assert len(np.unique(expect)) >= 20
    assert _eq(got, expect), name
    d = meta["view"]["depth"]
    assert st["pixel_iters"] == float(np.where(expect > 0, expect, d).sum())


def test_deep_maps_in_one_batch(ctx, golden):
    """All deep crops as the tiles of one batch (every arithmetic class at
    once): persistent and per-tile forms both equal the maps."""
    deep = [(m, e) for m, e in golden if m["name"].startswith("deep_")]
    specs = [(_golden_view(m["view"]), None, None) for m, _ in deep]
    for mode in (M.BATCH_PERSISTENT, M.BATCH_PER_TILE):
        print("deep batch mode", mode, flush=True)
        outs, _ = ctx.render_batch(specs, mode=mode)
        for (m, e), got in zip(deep, outs):
            assert _eq(got, e), (mode, m["name"])


def test_golden_maps(ctx, golden):
    """Every committed map (MDZ's own loops) reproduced on the device."""
    for meta, expect in golden:
        if meta["name"].startswith("deep_"):
            continue              # test_deep_golden_map
        vd = meta["view"]
        v = M.make_view(vd["rect"], precision=vd["precision"], depth=vd["depth"],
                        width=vd["width"], height=vd["height"], fractal=vd["fractal"],
                        family=vd["family"], path=vd["path"], julia=tuple(vd["julia"]))
        line0, nl, step = meta["lines"]
        col0, nc = meta["cols"]
        if step == 1:
            got, _ = ctx.render(v, lines=(line0, nl), cols=(col0, nc))
        else:
            full, _ = ctx.render(v, lines=(line0, (nl - 1) * step + 1), cols=(col0, nc))
            got = full[::step]
        print(meta["name"], flush=True)
        assert _eq(got, expect), meta["name"]


@pytest.mark.parametrize("name", sorted(cases.SCENES))
def test_gallery_scenes(ctx, oracle, name):
    rect, p, depth, fractal, path, (W, H) = cases.scene_rect(name)
    lines, cols = (H // 4, min(12, H // 2)), (W // 4, 40)
    v = M.make_view(rect, precision=p, depth=depth, width=W, height=H, fractal=fractal, path=path)
    got, _ = ctx.render(v, lines=lines, cols=cols)
    ref = oracle.render(rect, precision=p, depth=depth, width=W, height=H, fractal=fractal,
                        path=path, lines=(lines[0], lines[1], 1), cols=cols, threads=16)
    assert _eq(got, ref)


@pytest.mark.parametrize("name", sorted(cases.SCENES))
def test_scene_files(ctx, oracle, name):
    """A gallery .mdz file loaded and rendered as MDZ's command line renders it
    (its own path, its rect at 240 wide; corner-form files also refitted),
    against the oracle on that rect."""
    import os
    f = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mdz", name + ".mdz")
    for fit in ((False, True) if cases.SCENES[name][0] == "corner" else (False,)):
        v = M.scene_view(f, 240, fit_rect=fit)
        W, H = v.real_width, v.real_height
        lines, cols = (H // 3, min(6, H // 2)), (W // 4, 48)
        print(name, fit, v.path, W, H, flush=True)
        got, _ = ctx.render(v, lines=lines, cols=cols)
        ref = oracle.render(_rect(v), precision=v.precision, depth=v.depth, width=W, height=H,
                            fractal=v.fractal, path=v.path, lines=(lines[0], lines[1], 1),
                            cols=cols, threads=16)
        assert _eq(got, ref), (name, fit)


@pytest.mark.parametrize("p", [2, 17, 32, 33, 53, 63, 64, 65, 80, 96, 100, 127, 128, 129,
                               176, 184, 192, 255, 256, 320, 384, 500, 512, 640, 768, 1000, 1024])
def test_every_precision_class(ctx, oracle, p):
    """Word-aligned and ragged precisions of every word count, on a view with
    escape-time structure (the 'test' scene's centre at a shallower zoom)."""
    rect = oracle.rect_from_center("-1.7495931476527795538581440", "5.79069164258e-8",
                                   "3e-9", p, 96, 72)
    for fractal in (0, 2):
        v = M.make_view(rect, precision=p, depth=400, width=96, height=72, fractal=fractal)
        got, _ = ctx.render(v, lines=(30, 6), cols=(0, 96))
        ref = oracle.render(rect, precision=p, depth=400, width=96, height=72, fractal=fractal,
                            lines=(30, 6, 1), threads=16)
        assert _eq(got, ref), (p, fractal)


@pytest.mark.parametrize("p", [96, 100, 128, 200, 256, 500, 512, 1000, 1024])
def test_short_product_fallback(ctx, oracle, p):
    """Start values that make the first products exact put their guard word on
    the rounding boundary, where the short products must recompute in full:
    for ragged p dyadic c (guard word 0); for p = 32N, c = 1/2 + 2^(-1-16N),
    whose square is an exact tie (guard word 2^31) in column 0."""
    from fractions import Fraction as F
    N = p // 32
    if p == 32 * N:
        xmin, width, ymax, W, H = F(1, 2) + F(1, 2 ** (1 + 16 * N)), F(1, 16), F(1, 8), 16, 8
    else:
        xmin, width, ymax, W, H = F(-2), F(4), F(3, 2), 32, 24
    rect = [cases.hexs(xmin), cases.hexs(xmin + width), cases.hexs(ymax), cases.hexs(width)]
    for fractal in (0, 1):
        v = M.make_view(rect, precision=p, depth=200, width=W, height=H, fractal=fractal)
        got, _ = ctx.render(v)
        ref = oracle.render(rect, precision=p, depth=200, width=W, height=H, fractal=fractal,
                            threads=16)
        assert _eq(got, ref), (p, fractal)


# the GMP path: MDZ's truncating mpf loops (src/frac_mandel.c:55-80 and
# siblings, src/fractal.c:260-397); these scenes include ones where its counts
# differ from the MPFR path's (test, subdudehue, spiralblues)
@pytest.mark.parametrize("name", ["test", "subdudehue", "spiralblues", "floral", "space_pad",
                                  "honeytrace", "confirm"])
def test_gmp_path_scenes(ctx, oracle, name):
    rect, p, depth, fractal, _, (W, H) = cases.scene_rect(name)
    depth = min(depth, 3000)
    lines, cols = (H // 3, min(6, H // 2)), (0, W)
    v = M.make_view(rect, precision=p, depth=depth, width=W, height=H, fractal=fractal,
                    path=M.PATH_GMP)
    got, _ = ctx.render(v, lines=lines, cols=cols)
    ref = oracle.render(rect, precision=p, depth=depth, width=W, height=H, fractal=fractal,
                        path=oracle.PATH_GMP, lines=(lines[0], lines[1], 1), cols=cols,
                        threads=16)
    assert _eq(got, ref), name


@pytest.mark.parametrize("p", [2, 53, 64, 80, 128, 129, 192, 256, 257, 320, 512, 513, 640,
                               1000, 1024])
def test_gmp_path_every_precision(ctx, oracle, p):
    """Every limb precision (PREC + 1 = 3 .. 18 limbs) in all three storage
    buckets, centre-form rect at the coords precision."""
    rect = oracle.rect_from_center("-1.7495931476527795538581440", "5.79069164258e-8",
                                   "3e-9", p, 96, 72, coords=True)
    for fractal in (0, 3):
        v = M.make_view(rect, precision=p, depth=400, width=96, height=72, fractal=fractal,
                        path=M.PATH_GMP)
        got, _ = ctx.render(v, lines=(30, 4), cols=(0, 96))
        ref = oracle.render(rect, precision=p, depth=400, width=96, height=72, fractal=fractal,
                            path=oracle.PATH_GMP, lines=(30, 4, 1), threads=16)
        assert _eq(got, ref), (p, fractal)


def test_gmp_path_julia_and_batch(ctx, oracle):
    """Julia on the GMP path, and GMP tiles of all buckets mixed with the other
    arithmetics in one persistent batch and one launch per tile."""
    J = cases.JULIA
    tiles, refs = [], []
    for p, fr in ((80, 0), (200, 1), (600, 2), (1024, 3)):
        v = M.make_view(J["rect"], precision=p, depth=600, width=96, height=72, fractal=fr,
                        family=M.FAMILY_JULIA, path=M.PATH_GMP, julia=J["c"])
        tiles.append((v, (20, 6), (0, 96)))
        refs.append(oracle.render(J["rect"], precision=p, depth=600, width=96, height=72,
                                  fractal=fr, family=oracle.FAMILY_JULIA, path=oracle.PATH_GMP,
                                  julia=J["c"], lines=(20, 6, 1), threads=16))
    rect, p, depth, fractal, path, (W, H) = cases.scene_rect("subdudehue")
    for pth in (M.PATH_MPFR, M.PATH_GMP):
        v = M.make_view(rect, precision=p, depth=2000, width=W, height=H, fractal=fractal,
                        path=pth)
        tiles.append((v, (50, 5), (0, W)))
        refs.append(oracle.render(rect, precision=p, depth=2000, width=W, height=H,
                                  fractal=fractal, path=pth, lines=(50, 5, 1), threads=16))
    outs_p, _ = ctx.render_batch(tiles, M.BATCH_PERSISTENT)
    outs_t, _ = ctx.render_batch(tiles, M.BATCH_PER_TILE)
    for a, b, r in zip(outs_p, outs_t, refs):
        assert _eq(a, r) and _eq(b, r)


@pytest.mark.parametrize("fractal", [0, 1, 2, 3])
def test_julia_family(ctx, oracle, fractal):
    J = cases.JULIA
    for path, p in ((M.PATH_MPFR, 96), (M.PATH_MPFR, 200), (M.PATH_LD, 64), (M.PATH_DOUBLE, 64)):
        v = M.make_view(J["rect"], precision=p, depth=J["depth"], width=96, height=72,
                        fractal=fractal, family=M.FAMILY_JULIA, path=path, julia=J["c"])
        got, _ = ctx.render(v)
        ref = oracle.render(J["rect"], precision=p, depth=J["depth"], width=96, height=72,
                            fractal=fractal, family=M.FAMILY_JULIA, path=path, julia=J["c"],
                            threads=16)
        assert _eq(got, ref), (path, p)


def test_config2_full_frame(ctx, oracle):
    v = WL.config2_view()
    got, st = ctx.render(v)
    assert st["pixel_iters"] == 640 * 480 * 5000.0     # every pixel runs to depth
    ref = oracle.render(_rect(v), precision=128, depth=5000, width=640, height=480,
                        lines=(0, 3, 239), threads=16)
    assert _eq(got[::239], ref)


def test_config5_crop(ctx, oracle):
    v = WL.config5_view()
    got, _ = ctx.render(v, lines=(1000, 4), cols=(2000, 8))
    ref = oracle.render(_rect(v), precision=1024, depth=20000, width=3840, height=2160,
                        lines=(1000, 4, 1), cols=(2000, 8), threads=16)
    assert _eq(got, ref)


def test_batch_modes_agree_with_oracle(ctx, oracle):
    """Mixed tiles (precision, depth, formula, size) in one persistent launch per
    class, and one launch per tile: identical counts, equal to the oracle."""
    rng = random.Random(3)
    specs, refs = [], []
    for name in ("test", "subdudehue", "satellite", "honeytrace", "spiralblues", "confirm"):
        rect, p, depth, fractal, path, (W, H) = cases.scene_rect(name)
        nl, nc = rng.randint(3, 9), rng.randint(5, 40)
        l0, c0 = rng.randrange(H - nl), rng.randrange(W - nc)
        v = M.make_view(rect, precision=p, depth=min(depth, 4000), width=W, height=H,
                        fractal=fractal, path=path)
        specs.append((v, (l0, nl), (c0, nc)))
        refs.append(oracle.render(rect, precision=p, depth=min(depth, 4000), width=W, height=H,
                                  fractal=fractal, path=path, lines=(l0, nl, 1), cols=(c0, nc),
                                  threads=16))
    v1 = WL.config1_view(fractal=1)
    specs.append((v1, (10, 20), (0, 240)))
    refs.append(oracle.render(_rect(v1), precision=64, depth=1000, width=240, height=180,
                              fractal=1, path=2, lines=(10, 20, 1), threads=16))
    a, sa = ctx.render_batch(specs, M.BATCH_PERSISTENT)
    b, sb = ctx.render_batch(specs, M.BATCH_PER_TILE)
    # a synchronous render has no sentinel fill: one launch per tile / per class
    assert sb["launches"] == len(specs) and sa["launches"] <= len(specs)
    for x, y, r in zip(a, b, refs):
        assert _eq(x, y) and _eq(x, r)


def test_lazy_squares_with_refills(ctx, oracle):
    """The 1024-bit word-aligned loop keeps only an exponent bound for wim^2
    while it lies 32N+2 bits below wre^2 on every lane of a wave.  Tiles A
    are 64 one-row tiles k e-200 above the real axis (k = 1..64), from
    c = 0.2501 to 0.2597: every lane is in that state until its pixel
    escapes, after 29 to 311 iterations, and neighbours in a wave escape at
    different iterations.  With one block per CU (65,536 lanes, the A tiles'
    pixel count) and a refill every iteration, lanes that finish take pixels
    of tile B, off the axis, into waves whose other lanes still hold bounds,
    which then square those lanes exactly.  Every tile equals the oracle, in
    the persistent batch and one launch per tile."""
    specs, refs = [], []
    for k in range(64):
        rect = ("0.2501", "0.2597", "%de-200" % (k + 1), "0.0096")
        specs.append((M.make_view(rect, precision=1024, depth=2000, width=1024, height=1),
                      (0, 1), (0, 1024)))
        refs.append(oracle.render(rect, precision=1024, depth=2000, width=1024, height=1,
                                  threads=16))
    rect_b = oracle.rect_from_center("0.2801", "0.0102", "1e-3", 1024, 64, 64)
    specs.append((M.make_view(rect_b, precision=1024, depth=600, width=64, height=64),
                  (0, 64), (0, 64)))
    refs.append(oracle.render(rect_b, precision=1024, depth=600, width=64, height=64,
                              threads=16))
    assert refs[0].min() > 0 and len(np.unique(refs[0])) > 100   # every A pixel escapes
    kid = M.kernel_id(M.PATH_MPFR, 1024)
    keep = ctx.tune_get(kid)[:2]
    ctx.tune_set(kid, blocks_per_cu=1, chunk=1)
    try:
        a, _ = ctx.render_batch(specs, M.BATCH_PERSISTENT)
    finally:
        ctx.tune_set(kid, blocks_per_cu=keep[0], chunk=keep[1])
    b, _ = ctx.render_batch(specs, M.BATCH_PER_TILE)
    for x, y, r in zip(a, b, refs):
        assert _eq(x, r) and _eq(y, r)


@pytest.mark.parametrize("prec", [384, 512, 768])
def test_lazy_squares_other_word_counts(ctx, oracle, prec):
    """The exponent-bound squares at 12, 16 and 24 words: 64 one-row tiles
    k e-200 above the real axis (wim^2 some 1300 bits below wre^2, beyond
    32N + 2 at every word count here) and an off-axis tile, in one
    persistent launch with one block per CU and a refill every iteration, so
    lanes of the off-axis tile join waves whose other lanes hold bounds.
    Every tile equals the oracle, in both batch forms."""
    specs, refs = [], []
    for k in range(64):
        rect = ("0.2501", "0.2597", "%de-200" % (k + 1), "0.0096")
        specs.append((M.make_view(rect, precision=prec, depth=2000, width=1024, height=1),
                      (0, 1), (0, 1024)))
        refs.append(oracle.render(rect, precision=prec, depth=2000, width=1024, height=1,
                                  threads=16))
    rect_b = oracle.rect_from_center("0.2801", "0.0102", "1e-3", prec, 64, 64)
    specs.append((M.make_view(rect_b, precision=prec, depth=600, width=64, height=64),
                  (0, 64), (0, 64)))
    refs.append(oracle.render(rect_b, precision=prec, depth=600, width=64, height=64,
                              threads=16))
    assert refs[0].min() > 0 and len(np.unique(refs[0])) > 100
    kid = M.kernel_id(M.PATH_MPFR, prec)
    keep = ctx.tune_get(kid)[:2]
    ctx.tune_set(kid, blocks_per_cu=1, chunk=1)
    try:
        a, _ = ctx.render_batch(specs, M.BATCH_PERSISTENT)
    finally:
        ctx.tune_set(kid, blocks_per_cu=keep[0], chunk=keep[1])
    b, _ = ctx.render_batch(specs, M.BATCH_PER_TILE)
    for x, y, r in zip(a, b, refs):
        assert _eq(x, r) and _eq(y, r)


@pytest.mark.parametrize("chunk", [1, 3, 4, 5, 256])
def test_fp64_runs_mixed_depths(ctx, oracle, chunk):
    """FP64 persistent runs: one launch whose waves hold lanes of small tiles
    with depths 1 to 4097, all four formulas and a Julia tile, at run lengths
    that use the unrolled part of a run, its single steps, or both.  Every map
    equals the oracle and the one-launch-per-tile form."""
    rect = _rect(WL.config1_view())
    J = cases.JULIA
    rng = random.Random(11)
    specs, refs = [], []
    for fractal, depth in ((0, 1), (0, 2), (1, 3), (2, 5), (3, 7), (0, 100), (3, 257),
                           (1, 1000), (0, 4097)):
        nl, nc = 7, 13
        l0, c0 = rng.randrange(180 - nl), rng.randrange(240 - nc)
        v = M.make_view(rect, precision=64, depth=depth, width=240, height=180,
                        fractal=fractal, path=M.PATH_DOUBLE)
        specs.append((v, (l0, nl), (c0, nc)))
        refs.append(oracle.render(rect, precision=64, depth=depth, width=240, height=180,
                                  fractal=fractal, path=M.PATH_DOUBLE, lines=(l0, nl, 1),
                                  cols=(c0, nc), threads=8))
    vj = M.make_view(J["rect"], precision=64, depth=J["depth"], width=96, height=72, fractal=3,
                     family=M.FAMILY_JULIA, path=M.PATH_DOUBLE, julia=J["c"])
    specs.append((vj, (30, 9), (40, 11)))
    refs.append(oracle.render(J["rect"], precision=64, depth=J["depth"], width=96, height=72,
                              fractal=3, family=M.FAMILY_JULIA, path=M.PATH_DOUBLE,
                              julia=J["c"], lines=(30, 9, 1), cols=(40, 11), threads=8))
    kid = M.kernel_id(M.PATH_DOUBLE, 64)
    keep = ctx.tune_get(kid)[1]
    ctx.tune_set(kid, chunk=chunk)
    try:
        a, sa = ctx.render_batch(specs, M.BATCH_PERSISTENT)
        b, _ = ctx.render_batch(specs, M.BATCH_PER_TILE)
    finally:
        ctx.tune_set(kid, chunk=keep)
    assert sa["launches"] == 1
    for x, y, r in zip(a, b, refs):
        assert _eq(x, r) and _eq(y, r)
    assert sa["pixel_iters"] == sum(float(np.where(r > 0, r, s[0].depth).sum())
                                    for r, s in zip(refs, specs))


def test_tile_split_invariance(ctx):
    """A rectangle rendered whole equals the same rectangle rendered as four tiles."""
    rect, p, depth, fractal, path, (W, H) = cases.scene_rect("deep_embedded_julia")
    v = M.make_view(rect, precision=p, depth=3000, width=W, height=H)
    whole, _ = ctx.render(v, lines=(40, 40), cols=(60, 64))
    parts, _ = ctx.render_batch([(v, (40, 20), (60, 32)), (v, (40, 20), (92, 32)),
                                 (v, (60, 20), (60, 32)), (v, (60, 20), (92, 32))])
    top = np.hstack([parts[0], parts[1]])
    bot = np.hstack([parts[2], parts[3]])
    assert _eq(whole, np.vstack([top, bot]))


def test_frame_adapter_threads(ctx):
    """next_line_cb shape: worker threads claim rows; result equals a batch render."""
    rect, p, depth, fractal, path, (W, H) = cases.scene_rect("subdudehue")
    v = M.make_view(rect, precision=p, depth=2000, width=W, height=H)
    raw = np.full((H, W), -1, dtype=np.int32)
    fr = M.Frame(ctx, v, raw)
    nxt = [0]
    lock = threading.Lock()
    rets = []

    def worker():
        while True:
            with lock:
                line = nxt[0]
                nxt[0] += 1
            if line >= H:
                return
            rets.append(fr.calculate_line(line))

    ts = [threading.Thread(target=worker) for _ in range(8)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    rc, st = fr.close()
    assert rc == 0 and all(r == 1 for r in rets)
    whole, _ = ctx.render(v)
    assert _eq(raw, whole)


def test_frame_stop_keeps_sentinel(ctx):
    """A stop before the frame is done: later calls return 0 and rows never
    rendered keep the caller's sentinel (src/image_info.c:214-215)."""
    v = WL.config2_view()          # ~1.5e9 pixel-iterations, several bands
    big = M.make_view(_rect(v), precision=128, depth=5000, width=640, height=2048)
    raw = np.full((2048, 640), -1, dtype=np.int32)
    fr = M.Frame(ctx, big, raw)
    assert fr.calculate_line(0) == 1
    fr.stop()
    assert fr.calculate_line(2047) == 0
    rc, _ = fr.close()
    assert rc == -4
    assert (raw[0] == 0).all()
    assert (raw[-1] == -1).all()


def test_async_submit_and_cancel(ctx, golden):
    # a 1024-bit boundary view (escapes from ~3000 to 20000 iterations) at
    # 192 x 192, some 40 ms of device work, cancelled right after submission
    meta = next(m for m, _ in golden if m["name"] == "deep_mpfr1024_f2")
    vd = dict(meta["view"], width=192, height=192)
    cx, cy = meta["view"]["rect"][0], meta["view"]["rect"][2]   # corner of the crop
    rect = M.rect_from_center(cx, cy, "1e-249", 1024, 192, 192)
    v = M.make_view(rect, precision=1024, depth=vd["depth"], width=192, height=192,
                    fractal=2)
    full, _ = ctx.render(v)
    assert len(np.unique(full)) > 100
    SENT = 0xFFFFFFFF
    job = ctx.submit([(v, None, None)], fill=SENT)
    job.cancel()
    rc, st = job.wait()
    job.free()
    assert rc in (0, -4)
    c = job.counts[0]
    done = c != SENT
    # finished pixels equal the uncancelled render; unfinished keep the caller's
    # value (a job that was cancelled reports -4 whether or not the device had
    # finished every pixel when the flag landed)
    assert _eq(c[done], full[done])
    if rc == 0:
        assert done.all()
    job2 = ctx.submit([(WL.config1_view(), None, None)])
    while not job2.poll():
        pass
    rc2, st2 = job2.wait()
    job2.free()
    assert rc2 == 0 and st2["pixel_iters"] > 0


def test_rejects_bad_views_and_tiles(ctx):
    """Views and tiles outside what the ABI accepts are errors, not renders."""
    base = dict(precision=128, depth=100, width=16, height=12)
    rect = ("-2", "2", "1.5", "4")

    def code(rect=rect, lines=None, cols=None, **kw):
        args = dict(base)
        args.update(kw)
        with pytest.raises(M.MdzgError) as e:
            ctx.render(M.make_view(rect, **args), lines=lines, cols=cols)
        return e.value.code

    assert code(precision=1025) == -7
    assert code(precision=1) == -7
    assert code(precision=1025, path=M.PATH_GMP) == -7
    assert code(depth=0) == -1
    assert code(depth=2 ** 31) == -1
    assert code(fractal=4) == -1
    assert code(family=2) == -1
    assert code(path=4) == -1
    assert code(lines=(10, 3)) == -1                       # rows beyond the frame
    assert code(cols=(0, 17)) == -1
    assert code(rect=("x", "2", "1.5", "4")) == -6
    assert code(family=M.FAMILY_JULIA) == -6               # no Julia constant


@pytest.mark.parametrize("path", [M.PATH_MPFR, M.PATH_LD, M.PATH_DOUBLE, M.PATH_GMP])
def test_extreme_but_valid_views(ctx, oracle, path):
    """1x1 frames, depth 1, depth INT32_MAX on escaping pixels, one long row."""
    p = 2 if path in (M.PATH_MPFR, M.PATH_GMP) else 64
    v = M.make_view(("2", "3", "0", "1"), precision=p, depth=2 ** 31 - 1, width=1, height=1,
                    path=path)
    got, st = ctx.render(v)
    assert got.tolist() == [[1]] and st["pixel_iters"] == 1.0      # c = 2: |z2|^2 = 36 > 4
    v = M.make_view(("-2", "2", "1.5", "4"), precision=p, depth=1, width=40, height=30,
                    path=path)
    got, _ = ctx.render(v)
    ref = oracle.render(["-2", "2", "1.5", "4"], precision=p, depth=1, width=40, height=30,
                        path=path, threads=8)
    assert _eq(got, ref)
    W = 20000
    v = M.make_view(("-2", "1", "0.0001", "3"), precision=p, depth=60, width=W, height=3,
                    path=path)
    got, _ = ctx.render(v, lines=(1, 1))
    ref = oracle.render(["-2", "1", "0.0001", "3"], precision=p, depth=60, width=W, height=3,
                        path=path, lines=(1, 1, 1), threads=16)
    assert _eq(got, ref)


def test_config5_band_at_full_width(ctx, oracle):
    """Config 5 at its real row width (3840 samples, 1024 bits, depth 20000):
    a 16-row band rendered whole, as four column tiles in one persistent
    batch, and as four launches, gives one map; pixels spread over the band
    equal the oracle's."""
    v = WL.config5_view()
    l0, nl = 1048, 16
    whole, _ = ctx.render(v, lines=(l0, nl))
    specs = [(v, (l0, nl), (c, 960)) for c in range(0, 3840, 960)]
    parts, _ = ctx.render_batch(specs, M.BATCH_PERSISTENT)
    assert _eq(whole, np.hstack(parts))
    parts, _ = ctx.render_batch(specs, M.BATCH_PER_TILE)
    assert _eq(whole, np.hstack(parts))
    rng = random.Random(5)
    for _ in range(6):
        r, c = rng.randrange(nl), rng.randrange(3840)
        ref = oracle.render(_rect(v), precision=1024, depth=20000, width=3840, height=2160,
                            lines=(l0 + r, 1, 1), cols=(c, 1), threads=1)
        assert int(whole[r, c]) == int(ref[0, 0]), (r, c)


def test_config3_whole_batch_persistent_equals_per_tile(ctx):
    """Config 3's 256 tiles (every precision class, FP64 to 1024 bits, depths
    1e3 to 5e4, three formulas) in one persistent batch and as 256 launches:
    identical maps and pixel-iterations."""
    specs = [(WL.spec_view(sp), None, None) for sp in WL.config3_specs()]
    print("config 3 persistent", flush=True)
    a, sa = ctx.render_batch(specs, M.BATCH_PERSISTENT)
    print("config 3 per tile", flush=True)
    b, sb = ctx.render_batch(specs, M.BATCH_PER_TILE)
    for i, (x, y) in enumerate(zip(a, b)):
        assert _eq(x, y), i
    assert sa["pixel_iters"] == sb["pixel_iters"]
