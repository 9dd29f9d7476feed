"""GPU tests of the continuous-batching server: each request's counts are the
same alone, in a mix, across many slices (orbit state saved and resumed in
HBM), under a tight HBM budget, and in the per-request baseline form."""
import numpy as np
import pytest

import mdz_amd as M
from mdz_amd import workloads as WL
from tests import cases

pytestmark = pytest.mark.gpu


def _eq(a, b):
    return np.array_equal(np.asarray(a).astype(np.int64), np.asarray(b).astype(np.int64))


def _mix():
    """Requests across every arithmetic: MPFR at aligned and ragged p, long
    double, FP64, GMP, all four formulas, Julia."""
    reqs = []
    for name in ("test", "subdudehue", "satellite", "honeytrace", "spiralblues", "confirm",
                 "mdz_banner", "floral"):
        rect, p, depth, fractal, path, (W, H) = cases.scene_rect(name)
        v = M.make_view(rect, precision=p, depth=min(depth, 3000), width=W, height=H,
                        fractal=fractal, path=path)
        reqs.append((v, (H // 3, min(12, H // 3)), (W // 5, 48)))
    v1 = WL.config1_view(fractal=2)
    reqs.append((v1, (30, 40), (0, 240)))
    for name in ("test", "floral"):              # the GMP path (truncating mpf)
        rect, p, depth, fractal, _, (W, H) = cases.scene_rect(name)
        v = M.make_view(rect, precision=p, depth=min(depth, 2000), width=W, height=H,
                        fractal=fractal, path=M.PATH_GMP)
        reqs.append((v, (H // 3, 6), (W // 5, 40)))
    J = cases.JULIA
    for fr in range(4):
        vj = M.make_view(J["rect"], precision=96, depth=800, width=96, height=72, fractal=fr,
                         family=M.FAMILY_JULIA, julia=J["c"])
        reqs.append((vj, (10, 20), (0, 96)))
    return reqs


@pytest.mark.parametrize("slice_iters", [0, 7])
def test_server_results_equal_standalone(ctx, slice_iters):
    reqs = _mix()
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c in reqs]
    with M.Server(ctx, slice_iters=slice_iters) as srv:
        ids = [srv.submit(v, l, c) for v, l, c in reqs]
        outs = [srv.wait(i) for i in ids]
        st = srv.stats()
    assert st["completed"] == len(reqs)
    if slice_iters == 7:
        assert st["slices"] > 100           # many save/resume cycles per pixel
    for (out, info), ref in zip(outs, alone):
        assert info["status"] == 2 and info["latency_ms"] >= 0
        assert _eq(out, ref)


def test_server_unequal_requests_in_one_class(ctx):
    """Requests of one class whose work differs by far more than 8x: the
    slice caps each lane's iterations and lets pixels run on within the cap
    (the queue ordered by work left).  Counts equal standalone renders,
    across many slices, with the large request arriving first and last."""
    v = WL.config2_view()
    ve = WL.config2_escaping_view()
    big = (v, (200, 96), (300, 96))
    small = [(ve, (40 + 70 * i, 8), (100 + 90 * i, 8)) for i in range(4)]
    for reqs in ([big] + small, small + [big]):
        alone = [ctx.render(x, lines=l, cols=c)[0] for x, l, c in reqs]
        with M.Server(ctx, slice_iters=50) as srv:
            ids = [srv.submit(x, l, c) for x, l, c in reqs]
            outs = [srv.wait(i) for i in ids]
            assert srv.stats()["slices"] > 20
        for (out, info), ref in zip(outs, alone):
            assert info["status"] == 2
            assert _eq(out, ref)


def test_server_tight_hbm_budget(ctx):
    """Only a few requests fit at once: admission queues the rest, results hold."""
    reqs = _mix() * 2
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c in reqs]
    budget = 4 << 20
    with M.Server(ctx, hbm_budget=budget, slice_iters=64) as srv:
        ids = [srv.submit(v, l, c) for v, l, c in reqs]
        outs = [srv.wait(i)[0] for i in ids]
        st = srv.stats()
    assert st["hbm_peak"] <= budget
    for out, ref in zip(outs, alone):
        assert _eq(out, ref)


def test_server_per_request_mode(ctx):
    reqs = _mix()
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c in reqs]
    with M.Server(ctx, mode=M.SERVE_PER_REQUEST) as srv:
        ids = [srv.submit(v, l, c) for v, l, c in reqs]
        outs = [srv.wait(i)[0] for i in ids]
    for out, ref in zip(outs, alone):
        assert _eq(out, ref)


def test_server_rejects_request_larger_than_budget(ctx):
    v = WL.config2_view()
    with M.Server(ctx, hbm_budget=1 << 20) as srv:
        with pytest.raises(M.MdzgError) as e:
            srv.submit(v)
        assert e.value.code == -5


def test_server_arrivals_while_running(ctx):
    """Requests submitted while earlier ones are on the device join and finish."""
    v = WL.config2_view()
    import time
    with M.Server(ctx, slice_iters=256) as srv:
        first = srv.submit(v)                     # 1.5e9 pixel-iterations
        time.sleep(0.002)
        small = [srv.submit(v, (8 * k, 8), (0, 64)) for k in range(10)]
        outs = [srv.wait(i) for i in small]
        big, info_big = srv.wait(first)
    ref = ctx.render(v)[0]
    assert _eq(big, ref)
    for k, (o, info) in enumerate(outs):
        assert _eq(o, ref[8 * k:8 * k + 8, :64])
        assert info["admit_ms"] < info_big["done_ms"]     # admitted while the big one ran


@pytest.mark.parametrize("mode", [M.SERVE_BATCHED, M.SERVE_PER_REQUEST])
def test_server_image_requests(ctx, oracle, mode):
    """Image requests mixed with count requests: colours equal the oracle's
    colouring of the standalone counts (aa 1..3, interpolated or not), counts
    returned when asked, pixel-iterations as counted from the counts."""
    reqs = _mix()
    rng = np.random.default_rng(3)
    cols = [int(x) for x in rng.integers(0, 1 << 32, 256, dtype=np.uint64)]
    plans = []
    for k, (v, (l0, nl), (c0, nc)) in enumerate(reqs):
        aa = (1, 2, 3)[k % 3]
        nl, nc = nl - nl % aa, nc - nc % aa
        ip = bool(k % 2)
        pal = M.make_palette(cols[:64 + 12 * k], offset=(17 * k) % 256,
                             colour_scale=(0.37, 1.0, 0.0253907)[k % 3], interpolate=ip)
        plans.append((v, (l0, nl), (c0, nc), aa, pal, k % 4 == 0))
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c, _, _, _ in plans]
    with M.Server(ctx, mode=mode, slice_iters=9 if mode == M.SERVE_BATCHED else 0) as srv:
        ids = []
        for k, (v, l, c, aa, pal, plain) in enumerate(plans):
            if plain:
                ids.append(srv.submit(v, l, c))
            else:
                ids.append(srv.submit_image(v, aa, pal, l, c, counts=(k % 4 == 1)))
        outs = [srv.wait(i) for i in ids]
    for k, ((v, l, c, aa, pal, plain), ref, (out, info)) in enumerate(zip(plans, alone, outs)):
        assert info["status"] == 2
        assert info["pixel_iters"] == float(np.where(ref > 0, ref, v.depth).sum())
        if plain:
            assert _eq(out, ref)
            continue
        rgb = out[0] if k % 4 == 1 else out
        if k % 4 == 1:
            assert _eq(out[1], ref)
        want = oracle.colour(ref, aa, M.palette_colours(pal), offset=pal.offset,
                             colour_scale=pal.colour_scale, interpolate=bool(pal.interpolate))
        assert np.array_equal(rgb, want), k


def test_server_image_request_arguments(ctx):
    v = WL.config1_view()
    pal = M.make_palette([1, 2, 3])
    with M.Server(ctx) as srv:
        with pytest.raises(ValueError):
            srv.submit_image(v, 2, pal, (0, 5), (0, 8))      # 5 rows: not whole 2x2 blocks
        bad = M.make_palette([1, 2, 3])
        bad.n = 300
        with pytest.raises(M.MdzgError):
            srv.submit_image(v, 1, bad, (0, 4), (0, 4))


@pytest.mark.parametrize("graphs", [0, 1])
def test_server_slices_with_and_without_graphs(ctx, graphs):
    """Slices replayed as cached HIP graphs (the default) and launched one by
    one give the standalone counts; several replays of each graph shape run
    (slice_iters 7: many slices per request)."""
    reqs = _mix()
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c in reqs]
    ctx.option_set(M.OPT_GRAPHS, graphs)
    try:
        with M.Server(ctx, slice_iters=7) as srv:
            ids = [srv.submit(v, l, c) for v, l, c in reqs]
            outs = [srv.wait(i) for i in ids]
            assert srv.stats()["slices"] > 100
    finally:
        ctx.option_set(M.OPT_GRAPHS, 1)
    for (out, info), ref in zip(outs, alone):
        assert info["status"] == 2
        assert _eq(out, ref)


def test_server_staging_failure_fails_only_that_request(ctx, monkeypatch):
    """A request whose staging fails (forced from the host by the stage
    failpoint) retires with the error; the requests around it, staged in the
    same round and resident across its retirement, finish with their
    standalone counts."""
    reqs = _mix()
    alone = [ctx.render(v, lines=l, cols=c)[0] for v, l, c in reqs]
    monkeypatch.setenv("MDZG_FAILPOINT_STAGE", "3")      # the third request of this server
    with M.Server(ctx, slice_iters=7) as srv:
        ids = [srv.submit(v, l, c) for v, l, c in reqs]
        monkeypatch.delenv("MDZG_FAILPOINT_STAGE")
        infos = {}
        outs = {}
        for i in ids:
            try:
                outs[i], infos[i] = srv.wait(i)
            except M.MdzgError as e:
                infos[i] = e.code
    assert ids[2] == 3
    assert infos[3] == -3
    for i, ref in zip(ids, alone):
        if i == 3:
            continue
        assert infos[i]["status"] == 2, i
        assert _eq(outs[i], ref), i


def test_server_lazy_squares_across_slices(ctx, oracle):
    """1024-bit rows just above the real axis, where the word-aligned loop
    keeps only an exponent bound for wim^2, served in slices of 7 iterations
    next to an off-axis request of the same class: a slice that stops on a
    lane with a bound saves the exact square, and resumed lanes mix with
    lanes holding bounds.  Counts equal the oracle."""
    reqs, refs = [], []
    for k in range(8):
        rect = ("0.2501", "0.2597", "%de-200" % (k + 1), "0.0096")
        reqs.append((M.make_view(rect, precision=1024, depth=2000, width=1024, height=1),
                     (0, 1), (0, 1024)))
        refs.append(oracle.render(rect, precision=1024, depth=2000, width=1024, height=1,
                                  threads=16))
    rect_b = oracle.rect_from_center("0.2801", "0.0102", "1e-3", 1024, 64, 64)
    reqs.append((M.make_view(rect_b, precision=1024, depth=600, width=64, height=64),
                 (0, 16), (0, 64)))
    refs.append(oracle.render(rect_b, precision=1024, depth=600, width=64, height=64,
                              lines=(0, 16, 1), threads=16))
    with M.Server(ctx, slice_iters=7) as srv:
        ids = [srv.submit(v, l, c) for v, l, c in reqs]
        outs = [srv.wait(i) for i in ids]
        assert srv.stats()["slices"] > 20
    for (out, info), ref in zip(outs, refs):
        assert info["status"] == 2
        assert _eq(out, ref)
