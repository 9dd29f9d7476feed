"""Scene files (.mdz settings sections): the loader against the gallery files'
known contents, and the rect a scene is rendered with against an exact
restatement of MDZ's coords chain.  The fixtures under tests/golden/mdz are
the settings sections of MDZ's gallery/*.mdz files (palette sections cut)."""
import glob
import os
from fractions import Fraction

import pytest

from tests import cases, rn_ref

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MDZ = os.path.join(ROOT, "tests", "golden", "mdz")

# path each gallery file selects through its multi-precision / rounding flags
# (src/image_info.c:616-621 -> src/fractal.c:29-64)
PATHS = {"spiralblues": 1, "floral": 3, "space_pad": 3}


def _files():
    return sorted(os.path.basename(f)[:-4] for f in glob.glob(os.path.join(MDZ, "*.mdz")))


def test_every_gallery_scene_has_a_fixture():
    assert set(cases.SCENES) <= set(_files())


@pytest.mark.parametrize("name", sorted(cases.SCENES))
def test_scene_fields(name):
    import mdz_amd as M
    d = M.load_scene(os.path.join(MDZ, name + ".mdz"))
    form, p, depth, a, b, c, fractal, _, (w, h) = cases.SCENES[name]
    assert d["precision"] == p and d["depth"] == depth and d["fractal"] == fractal
    assert d["family"] == M.FAMILY_MANDEL and d["julia_re"] is None
    assert d["centre"] == (form == "centre")
    assert (d["c0"], d["c1"], d["c2"]) == (a, b, c)
    assert d["path"] == PATHS.get(name, M.PATH_MPFR)
    assert abs(d["aspect"] - w / h) < 1e-2 * (w / h)     # the parity views round it


@pytest.mark.parametrize("name", sorted(cases.SCENES))
@pytest.mark.parametrize("wh", [(0, 0), (240, 0), (100, 300)])
@pytest.mark.parametrize("fit", [False, True])
def test_scene_rect_exact(name, wh, fit):
    """Centre form: coords_to + coords_size then coords_center_to_rect;
    corner form: the command line's chain (or the refitted rect) -- every step
    one RN at the coords precision."""
    import mdz_amd as M
    f = os.path.join(MDZ, name + ".mdz")
    d = M.load_scene(f)
    v = M.scene_view(f, *wh, fit_rect=fit)
    W, H = v.real_width, v.real_height
    h0 = int(480 / d["aspect"])
    if wh == (0, 0):
        assert (W, H) == (480, h0)
    elif wh[1] == 0:
        assert (W, H) == (wh[0], int(wh[0] / (480 / h0)))
    got = [rn_ref.parse(s.decode()) for s in v._keep[:4]]
    gmp = d["path"] == M.PATH_GMP
    args = (d["c0"], d["c1"], d["c2"], d["precision"], W, H)
    if d["centre"]:
        want = rn_ref.rect_from_center(*args, coords=gmp)
    else:
        want = rn_ref.rect_from_corners(*args, coords=gmp,
                                        file_aspect=None if fit else d["aspect"])
    assert got == list(want)
    assert (v.precision, v.depth, v.fractal, v.path) == \
        (d["precision"], d["depth"], d["fractal"], d["path"])


CORNER = [n for n in sorted(cases.SCENES) if cases.SCENES[n][0] == "corner"]


def _ulp16(d):
    f = [rn_ref.parse(s) for s in (d["c0"], d["c1"], d["c2"])]
    mag = max(abs(f[0]), abs(f[1]))
    return f, 16 * Fraction(2) ** (rn_ref.ilog2(mag) + 1 - max(d["precision"], 80))


@pytest.mark.parametrize("name", CORNER)
def test_corner_scene_on_the_command_line(name):
    """MDZ's command line renders a corner-form file at its rect's centre but
    at coords' stale size: the width (wide output) or height (tall) is exactly
    4.0, and the centre is the file rect's centre at the 480-wide load aspect."""
    import mdz_amd as M
    f = os.path.join(MDZ, name + ".mdz")
    d = M.load_scene(f)
    fv, ulp = _ulp16(d)
    for W, H in ((240, 180), (90, 300)):
        v = M.scene_view(f, W, H)
        xmin, xmax, ymax, width = (rn_ref.parse(s.decode()) for s in v._keep[:4])
        aspect = Fraction(W / H)
        if W > H:
            assert width == 4
        else:
            assert abs(width - 4 * aspect) <= 4 * Fraction(2) ** -70
        assert abs((xmin + xmax) / 2 - (fv[0] + fv[1]) / 2) <= ulp
        h0 = int(480 / d["aspect"])
        cy = fv[2] - (fv[1] - fv[0]) / Fraction(480 / h0) / 2
        height = width / aspect
        assert abs((ymax - height / 2) - cy) <= 16 * ulp


@pytest.mark.parametrize("name", CORNER)
def test_corner_scene_fit_keeps_its_rect(name):
    """Refitted at (nearly) the file's own aspect, a corner-form file's rect
    comes back to a few units in the last place of its precision."""
    import mdz_amd as M
    f = os.path.join(MDZ, name + ".mdz")
    d = M.load_scene(f)
    W = 240
    H = int(W / d["aspect"])
    v = M.scene_view(f, W, H, fit_rect=True)
    xmin, xmax, ymax, width = (rn_ref.parse(s.decode()) for s in v._keep[:4])
    fv, ulp = _ulp16(d)
    span = fv[1] - fv[0]
    assert abs(xmin - fv[0]) <= ulp and abs(xmax - fv[1]) <= ulp
    assert abs(width - span) <= ulp
    assert abs(ymax - fv[2]) <= span * Fraction(1, 100)


def test_aa_multiplies_the_render_size():
    import mdz_amd as M
    v1 = M.scene_view(os.path.join(MDZ, "test.mdz"), 240, 180)
    v3 = M.scene_view(os.path.join(MDZ, "test.mdz"), 240, 180, aa=3)
    assert (v3.real_width, v3.real_height) == (720, 540)
    assert v3._keep[:4] == v1._keep[:4]


def test_julia_and_old_format(tmp_path):
    """A Julia file in the current format (the 0.0.x format is the 'test'
    gallery file's)."""
    import mdz_amd as M
    j = tmp_path / "j.mdz"
    j.write_text("mdz fractal settings 0.1.3\n# a comment\nsettings\n"
                 "family julia\nfractal burning ship\ndepth 2000\naspect\t1.3333333333\n"
                 "colour-scale 1.0\ncolour-interpolate yes\nmulti-precision yes\n"
                 "multi-rounding yes\nprecision 96\ncx 0.0\ncy 0.0\nsize 3.2\n"
                 "julia-real -0.8\njulia-imag 0.156\n")
    d = M.load_scene(str(j))
    assert (d["family"], d["fractal"], d["path"], d["precision"]) == (M.FAMILY_JULIA, 1,
                                                                       M.PATH_MPFR, 96)
    assert (d["julia_re"], d["julia_im"], d["centre"]) == ("-0.8", "0.156", 1)
    v = M.scene_view(str(j), 240)
    assert (v.real_width, v.real_height, v.family) == (240, 180, M.FAMILY_JULIA)
    assert [rn_ref.parse(s.decode()) for s in v._keep[:4]] == \
        list(rn_ref.rect_from_center("0.0", "0.0", "3.2", 96, 240, 180))
    assert [rn_ref.parse(s.decode()) for s in v._keep[4:6]] == [rn_ref.parse("-0.8"),
                                                                rn_ref.parse("0.156")]


def test_bad_files_are_errors(tmp_path):
    import mdz_amd as M
    bad = {
        "header": "not a settings file\n",
        "truncated": "mdz fractal settings 0.1.3\nsettings\nfamily mandelbrot\n",
        "value": "mdz fractal settings 0.1.3\nsettings\nfamily mandelbrot\nfractal cubic\n",
        "precision": "mdz fractal settings\nsettings\nfractal mandelbrot\ndepth 10\n"
                     "aspect 1.0\ncolour-scale 1.0\ncolour-interpolate no\nmpfr yes\n"
                     "precision 64\ncx 0\ncy 0\nsize 1\n",
    }
    for k, text in bad.items():
        f = tmp_path / (k + ".mdz")
        f.write_text(text)
        with pytest.raises(M.MdzgError):
            M.load_scene(str(f))
    with pytest.raises(M.MdzgError):
        M.load_scene(str(tmp_path / "missing.mdz"))
