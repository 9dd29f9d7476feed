"""GPU tests of the measurement entry points: the issue-rate probe and the
per-operation probe of the RN arithmetic.  They return times and rates, not
counts, so the tests check that each runs, returns a positive figure, and
rejects what it does not support."""
import pytest

import mdz_amd as M
from mdz_amd import workloads as WL

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("kind", [0, 1])
def test_rate_probe_positive(ctx, kind):
    v = ctx.rate_probe(kind, 2)
    assert v > 1e11


def test_rate_probe_rejects_bad_arguments(ctx):
    with pytest.raises(M.MdzgError):
        ctx.rate_probe(2, 2)
    with pytest.raises(M.MdzgError):
        ctx.rate_probe(0, 0)


@pytest.mark.parametrize("leg", ["rn128", "rn1024"])
def test_op_probe_every_operation(ctx, leg):
    view = WL.config2_view() if leg == "rn128" else WL.config5_view()
    lines = (0, 2) if leg == "rn128" else (1048, 1)
    for op in range(7):
        ms = ctx.op_probe(view, op, warm=20, reps=20, lines=lines, cols=(0, 256))
        assert ms > 0.0, (leg, op)


def test_op_probe_rejects_other_arithmetics(ctx):
    rect = M.rect_from_center("-0.75", "0", "1e-20", 256, 64, 64)
    v = M.make_view(rect, precision=256, depth=100, width=64, height=64)
    with pytest.raises(M.MdzgError):
        ctx.op_probe(v, 1, warm=1, reps=1)
    with pytest.raises(M.MdzgError):
        ctx.op_probe(WL.config2_view(), 7, warm=1, reps=1, lines=(0, 1))
