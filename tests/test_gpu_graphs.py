"""GPU tests of graph replay, the options and the autotuner.

A persistent batch of two or more arithmetic classes (the sentinel fill of
an asynchronous job, the concurrent class launches) runs as one replay of a
cached HIP graph by default; a one-class plan is launched directly.  Each
test replays graphs several times, with eager (one-by-one) renders and host
synchronisation between the replays, and checks every replay bit for bit
against the maps of MDZ's own loops."""
import os

import numpy as np
import pytest

import mdz_amd as M
from mdz_amd import workloads as WL

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return np.array_equal(np.asarray(a).astype(np.int64), np.asarray(b).astype(np.int64))


def _view(vd):
    return M.make_view(vd["rect"], precision=vd["precision"], depth=vd["depth"],
                       width=vd["width"], height=vd["height"], fractal=vd["fractal"],
                       family=vd["family"], path=vd["path"], julia=tuple(vd["julia"]))


def _deep(golden, names):
    return [(m, e) for m, e in golden if m["name"] in names]


MIXED = ["deep_double64_f0", "deep_ld64_f1", "deep_mpfr96_f2", "deep_mpfr128_f3",
         "deep_mpfr416_f0", "deep_mpfr1024_f2", "deep_gmp128_f1"]


def test_batch_graph_replays_equal_eager_and_golden(ctx, golden):
    cases = _deep(golden, MIXED)
    specs = [(_view(m["view"]), None, None) for m, _ in cases]
    for rep in range(4):
        for graphs in (1, 0):
            ctx.option_set(M.OPT_GRAPHS, graphs)
            print("replay", rep, "graphs", graphs, flush=True)
            outs, st = ctx.render_batch(specs)
            assert st["launches"] == len(MIXED)          # one launch per class
            for (m, e), got in zip(cases, outs):
                assert _eq(got, e), (rep, graphs, m["name"])
    ctx.option_set(M.OPT_GRAPHS, 1)


def test_single_tile_and_image_jobs_with_either_option(ctx, golden):
    """One-class batches and image jobs (kernel, colouring pass), launched
    directly under either setting of the option."""
    (m, e), = _deep(golden, ["deep_mpfr128_f0"])
    v = _view(m["view"])
    rng = np.random.default_rng(3)
    pal = M.make_palette([int(x) for x in rng.integers(0, 1 << 24, 64)], offset=5,
                         colour_scale=0.3, interpolate=True)
    ref_rgb = None
    for rep in range(3):
        for graphs in (1, 0):
            ctx.option_set(M.OPT_GRAPHS, graphs)
            got, _ = ctx.render(v)
            assert _eq(got, e)
            rgb, cnt, _ = ctx.render_image(v, 2, pal, counts=True)
            assert _eq(cnt, e)
            if ref_rgb is None:
                ref_rgb = rgb
            assert _eq(rgb, ref_rgb)
    ctx.option_set(M.OPT_GRAPHS, 1)
    assert _eq(ref_rgb, ctx.colour(e, 2, pal))


def test_options_validate(ctx):
    assert ctx.option_get(M.OPT_GRAPHS) == 1
    for opt, bad in ((M.OPT_GRAPHS, 2), (M.OPT_BAND_PIXELS, 0), (M.OPT_BAND_INFLIGHT, 17),
                     (99, 1)):
        with pytest.raises(M.MdzgError):
            ctx.option_set(opt, bad)
    b0 = ctx.option_get(M.OPT_BAND_PIXELS)
    ctx.option_set(M.OPT_BAND_PIXELS, 4096)
    assert ctx.option_get(M.OPT_BAND_PIXELS) == 4096
    ctx.option_set(M.OPT_BAND_PIXELS, b0)


def test_frame_adapter_with_small_bands(ctx, golden):
    """The line adapter with tiny bands (many jobs, many graph replays) and 1
    or 5 bands in flight equals the map."""
    (m, e), = _deep(golden, ["deep_mpfr256_f3"])
    v = _view(m["view"])
    b0 = ctx.option_get(M.OPT_BAND_PIXELS)
    i0 = ctx.option_get(M.OPT_BAND_INFLIGHT)
    try:
        for bp, inf in ((64, 1), (100, 5)):
            ctx.option_set(M.OPT_BAND_PIXELS, bp)
            ctx.option_set(M.OPT_BAND_INFLIGHT, inf)
            raw = np.full((v.real_height, v.real_width), -1, dtype=np.int32)
            fr = M.Frame(ctx, v, raw)
            for line in range(v.real_height):
                assert fr.calculate_line(line) == 1
            rc, _ = fr.close()
            assert rc == 0 and _eq(raw, e)
    finally:
        ctx.option_set(M.OPT_BAND_PIXELS, b0)
        ctx.option_set(M.OPT_BAND_INFLIGHT, i0)


def test_autotune_keeps_a_valid_faster_pair(ctx, golden, tmp_path):
    (m, e), = _deep(golden, ["deep_mpfr192_f1"])
    v = _view(m["view"])
    kid = M.kernel_id(M.PATH_MPFR, 192)
    before = ctx.tune_get(kid)
    r = ctx.autotune(v, reps=1)
    assert r["kernel_id"] == kid
    assert 1 <= r["blocks_per_cu"] <= before[2]
    assert r["chunk"] in (1, 2, 4, 8, 16, 32)
    assert r["candidates"] == before[2] * 6
    assert r["best_ms"] <= r["default_ms"]
    assert ctx.tune_get(kid)[:2] == (r["blocks_per_cu"], r["chunk"])
    got, _ = ctx.render(v)                 # tuned: the same counts
    assert _eq(got, e)
    # the table round-trips through a file
    path = str(tmp_path / "tune.txt")
    ctx.tune_save(path)
    ctx.tune_set(kid, blocks_per_cu=1, chunk=3)
    ctx.tune_load(path)
    assert ctx.tune_get(kid)[:2] == (r["blocks_per_cu"], r["chunk"])
    ctx.tune_set(kid, blocks_per_cu=before[0], chunk=before[1])
    bad = tmp_path / "bad.txt"
    bad.write_text("kernel 999 1 1\n")
    with pytest.raises(M.MdzgError):
        ctx.tune_load(str(bad))


def test_shipped_table_loads(ctx):
    """The tuned table beside the library parses and is what a new context holds."""
    path = os.path.join(os.path.dirname(M.LIB_PATH), "tune_gfx950.txt")
    if not os.path.exists(path):
        pytest.skip("no tuned table shipped beside the library")
    want = {}
    with open(path) as f:
        for line in f:
            p = line.split()
            if p and p[0] == "kernel" and int(p[2]) > 0 and int(p[3]) > 0:
                want[int(p[1])] = (int(p[2]), int(p[3]))
    with M.Context(0) as c2:
        for kid, (b, ch) in want.items():
            got = c2.tune_get(kid)
            assert got[1] == ch and got[0] == min(b, got[2]), kid


def test_async_jobs_fill_the_sentinel_in_the_graph(ctx, golden):
    """An asynchronous batch's graph starts with the sentinel fill: a job
    waited for without a cancel equals the maps, replay after replay."""
    cases = _deep(golden, MIXED)
    for rep in range(3):
        for graphs in (1, 0):
            ctx.option_set(M.OPT_GRAPHS, graphs)
            job = ctx.submit([(_view(m["view"]), None, None) for m, _ in cases],
                             fill=0x12345678)
            rc, st = job.wait()
            job.free()
            assert rc == 0 and st["launches"] == 1 + len(MIXED)
            for (m, e), got in zip(cases, job.counts):
                assert _eq(got, e), (rep, graphs, m["name"])
    ctx.option_set(M.OPT_GRAPHS, 1)


def test_failed_launch_releases_everything(golden, monkeypatch):
    """A job whose launch plan fails (forced from the host by the plan
    failpoint, after its copy-in was queued) returns the device error and
    leaves the context whole: the renders before and after it, and enough
    renders to cycle every cached buffer, equal the maps."""
    (m, e), = _deep(golden, ["deep_mpfr176_f3"])
    v = _view(m["view"])
    monkeypatch.setenv("MDZG_FAILPOINT_PLAN", "3")
    with M.Context(0) as c2:
        monkeypatch.delenv("MDZG_FAILPOINT_PLAN")
        for i in range(1, 13):
            if i == 3:
                with pytest.raises(M.MdzgError) as err:
                    c2.render(v)
                assert err.value.code == -3
                continue
            got, _ = c2.render(v)
            assert _eq(got, e), i
        job = c2.submit([(v, None, None)], fill=7)
        rc, _ = job.wait()
        job.free()
        assert rc == 0 and _eq(job.counts[0], e)
