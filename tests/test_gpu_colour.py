"""GPU colouring: the colour kernel (mdzg_colour) and whole-frame image jobs
(mdzg_render_image) against the oracle's restatement of MDZ's palette_apply /
do_anti_aliasing.  Tolerance: none (uint32 rgb_data must be bit-identical)."""
import os

import numpy as np
import pytest

import mdz_amd as M
from tests import cases

pytestmark = pytest.mark.gpu

MDZ = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mdz")


def _pal(n, seed):
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.integers(0, 1 << 32, n, dtype=np.uint64)]


def _counts(shape, seed, hi):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, hi, shape, dtype=np.int64).astype(np.uint32)
    c[rng.random(shape) < 0.2] = 0
    c.flat[:4] = [1, hi - 1, 0, 2]
    return c


@pytest.mark.parametrize("aa", [1, 2, 3, 4])
@pytest.mark.parametrize("ip", [False, True])
@pytest.mark.parametrize("n,offset,sc,hi", [(256, 0, 1.0, 100000),
                                            (256, 148, 0.02539073500000000108, 1 << 31),
                                            (37, 12, 0.18554686500000000526, 200000),
                                            (2, 1, 3.7, 1000), (3, 255, 1e-3, 1 << 31),
                                            (200, 7, 1048576.0, 1 << 31)])
def test_colour_kernel(ctx, oracle, aa, ip, n, offset, sc, hi):
    """Random counts (zeros, up to 2^31 - 1), palettes with bits above blue,
    every aa; scales whose values pass 2^32 exercise the index wrap."""
    if ip and aa == 1 and hi * sc + offset >= 2 ** 31:
        hi = int((2 ** 31 - 1 - offset) / sc)     # MDZ's own path is undefined above
    pal = _pal(n, n + aa)
    c = _counts((24 * aa, 40 * aa), aa * 13 + n, hi)
    p = M.make_palette(pal, offset=offset, colour_scale=sc, interpolate=ip)
    got = ctx.colour(c, aa, p)
    want = oracle.colour(c, aa, pal, offset=offset, colour_scale=sc, interpolate=ip)
    assert np.array_equal(got, want), (aa, ip, n, offset, sc)


def test_colour_kernel_odd_sizes(ctx, oracle):
    """Frames that are not a whole number of blocks of 256 pixels."""
    pal = _pal(256, 5)
    for (w, h, aa) in [(1, 1, 1), (1, 1, 5), (17, 3, 2), (257, 1, 1), (3, 255, 3)]:
        c = _counts((h * aa, w * aa), w * h, 50000)
        p = M.make_palette(pal, offset=5, colour_scale=0.37, interpolate=True)
        assert np.array_equal(ctx.colour(c, aa, p),
                              oracle.colour(c, aa, pal, offset=5, colour_scale=0.37,
                                            interpolate=True)), (w, h, aa)


@pytest.mark.parametrize("name,aa", [("floral", 1), ("honeytrace", 2), ("test", 1),
                                     ("satellite", 2), ("space_pad", 1)])
def test_render_image_gallery(ctx, oracle, name, aa):
    """A gallery file rendered and coloured on the device in one job, as
    `mdz -s file -w 96 -a aa` computes it: equal to the oracle's colouring of
    the counts, with the counts and the pixel-iteration count as a plain
    render gives them."""
    f = os.path.join(MDZ, name + ".mdz")
    v = M.scene_view(f, 96, aa=aa)
    v.depth = min(v.depth, 4000)
    pal = M.scene_palette(f)
    rgb, cnt, st = ctx.render_image(v, aa, pal, counts=True)
    ref_counts, st2 = ctx.render(v)
    assert np.array_equal(cnt, ref_counts)
    assert st["pixel_iters"] == st2["pixel_iters"] == float(
        np.where(ref_counts > 0, ref_counts, v.depth).sum())
    want = oracle.colour(ref_counts, aa, M.palette_colours(pal), offset=pal.offset,
                         colour_scale=pal.colour_scale, interpolate=bool(pal.interpolate))
    assert np.array_equal(rgb, want), name
    rgb2, st3 = ctx.render_image(v, aa, pal)       # counts left in device memory
    assert np.array_equal(rgb2, rgb) and st3["pixel_iters"] == st["pixel_iters"]


def test_render_image_to_png(ctx, tmp_path):
    from tests.test_colour_cpu import png_decode, _rgb_bytes
    f = os.path.join(MDZ, "mdz_banner.mdz")
    v = M.scene_view(f, 120, aa=2)
    pal = M.scene_palette(f)
    rgb, _ = ctx.render_image(v, 2, pal)
    out = tmp_path / "banner.png"
    M.png_write(str(out), rgb)
    w, h, img = png_decode(str(out))
    assert (w, h) == (rgb.shape[1], rgb.shape[0])
    assert np.array_equal(img, _rgb_bytes(rgb))


def test_render_image_fp64_and_julia(ctx, oracle):
    from mdz_amd import workloads as WL
    v = WL.config1_view()
    pal = M.make_palette(_pal(64, 9), offset=3, colour_scale=0.5, interpolate=True)
    rgb, cnt, st = ctx.render_image(v, 1, pal, counts=True)
    assert np.array_equal(rgb, oracle.colour(cnt, 1, _pal(64, 9), offset=3, colour_scale=0.5,
                                             interpolate=True))
    J = cases.JULIA
    vj = M.make_view(J["rect"], precision=J["precision"], depth=J["depth"], width=128, height=96,
                     family=M.FAMILY_JULIA, julia=J["c"])
    rgb, cnt, _ = ctx.render_image(vj, 2, pal, counts=True)
    ref, _ = ctx.render(vj)
    assert np.array_equal(cnt, ref)
    assert np.array_equal(rgb, oracle.colour(ref, 2, _pal(64, 9), offset=3, colour_scale=0.5,
                                             interpolate=True))


def test_render_image_rejects_bad_arguments(ctx):
    from mdz_amd import workloads as WL
    v = WL.config1_view()                          # 240 x 180
    good = M.make_palette([1, 2, 3])
    with pytest.raises(M.MdzgError):
        ctx.render_image(v, 7, good)               # not a whole number of 7x7 blocks
    bad = M.make_palette([1, 2, 3])
    bad.n = 1
    with pytest.raises(M.MdzgError):
        ctx.render_image(v, 1, bad)
    with pytest.raises(M.MdzgError):
        ctx.colour(np.zeros((4, 4), dtype=np.uint32), 2, bad)


def test_c_program_renders_a_scene_to_png(ctx, tmp_path):
    """tests/capi/render_mdz.c, the C program of INTEGRATION's standalone
    example, run as its own process: its PNG equals the image the Python
    binding renders for the same file."""
    import subprocess
    from tests.test_colour_cpu import png_decode, _rgb_bytes
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "capi", "render_mdz")
    if not os.path.exists(exe):
        from tests.capi import build as B
        B.build()
    f = os.path.join(MDZ, "honeytrace.mdz")
    out = tmp_path / "h.png"
    print("render_mdz honeytrace 160 aa 2", flush=True)
    r = subprocess.run([exe, f, "160", str(out), "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    v = M.scene_view(f, 160, aa=2)
    rgb, _ = ctx.render_image(v, 2, M.scene_palette(f))
    w, h, img = png_decode(str(out))
    assert (w, h) == (rgb.shape[1], rgb.shape[0])
    assert np.array_equal(img, _rgb_bytes(rgb))


def test_c_line_shim_under_a_row_claiming_pool(ctx, golden, tmp_path):
    """mdz_amd/shim/fractal_gpu.c, the int (*)(image_info*, int) callback MDZ
    installs (src/image_info.c:242-248), compiled as C against test-side
    image_info declarations and called by 16 pthreads that claim rows as
    src/render_threads.c:342-393 does: a full render equals the device
    render of the same view; a second render in the same process, stopped
    after 40 claims, leaves every row either complete and equal or at the
    UINT_MAX sentinel, as MDZ's stop does (src/image_info.c:214-215)."""
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = os.path.join(root, "tests", "capi", "shim_pool")
    if not os.path.exists(exe):
        from tests.capi import build as B
        B.build()
    meta = next(m for m, _ in golden if m["name"] == "deep_mpfr128_f2")
    vd = meta["view"]
    cx, cy = vd["rect"][0], vd["rect"][2]
    W, H = 640, 1536                       # four bands of the adapter's default size
    rect = M.rect_from_center(cx, cy, "1e-24", 128, W, H)
    v = M.make_view(rect, precision=128, depth=5000, width=W, height=H, fractal=2)
    want, _ = ctx.render(v)
    assert len(np.unique(want)) > 50
    out = tmp_path / "maps.bin"
    print("shim_pool 640x1536 p=128 celtic", flush=True)
    r = subprocess.run([exe, "0", "2", str(W), str(H), "128", "5000", *rect, "16", "2", "40",
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    done1, done2 = (int(x) for x in r.stdout.split()[1::2])
    maps = np.fromfile(str(out), dtype=np.uint32).reshape(2, H, W)
    assert done1 == H and np.array_equal(maps[0], want)
    rows = maps[1]
    full = (rows == want).all(axis=1)
    blank = (rows == 0xFFFFFFFF).all(axis=1)
    assert (full | blank).all()
    assert blank.sum() > 0 and full.sum() == done2 <= 2 * 40
