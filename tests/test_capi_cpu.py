"""The C ABI as a C compiler sees include/mdzg.h: a C program (tests/capi/
abi_layout.c) linked against libmdzg.so prints struct layouts, enum values
and the results of the calls that need no device; they must equal what the
ctypes mirror of mdz_amd sees and returns."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MDZ = os.path.join(ROOT, "tests", "golden", "mdz")


@pytest.fixture(scope="module")
def abi(tmp_path_factory):
    import mdz_amd  # noqa: F401  (the library must be built)
    d = tmp_path_factory.mktemp("capi")
    exe = d / "abi_layout"
    lib = os.path.join(ROOT, "mdz_amd")
    subprocess.run(["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-I",
                    os.path.join(ROOT, "include"), "-o", str(exe),
                    os.path.join(ROOT, "tests", "capi", "abi_layout.c"), "-L", lib, "-lmdzg",
                    "-Wl,-rpath," + lib], check=True)
    png = d / "t.png"
    r = subprocess.run([str(exe), os.path.join(MDZ, "floral.mdz"), str(png)],
                       capture_output=True, text=True, timeout=60, check=True)
    out = {}
    for line in r.stdout.splitlines():
        k, _, v = line.partition(" ")
        out[k] = v
    return out, png


def test_struct_layouts(abi):
    import mdz_amd as M
    out, _ = abi
    for name, cls in (("mdzg_view", M.View), ("mdzg_tile", M.Tile), ("mdzg_stats", M.Stats),
                      ("mdzg_palette", M.Palette), ("mdzg_scene", M.Scene),
                      ("mdzg_server_cfg", M.ServerCfg), ("mdzg_req_info", M.ReqInfo),
                      ("mdzg_serve_stats", M.ServeStats)):
        assert int(out["sizeof." + name]) == ctypes.sizeof(cls), name
    fields = [k for k in out if k.count(".") == 1 and k.startswith("mdzg_")]
    assert len(fields) >= 15
    cls = {"mdzg_view": M.View, "mdzg_tile": M.Tile, "mdzg_palette": M.Palette,
           "mdzg_scene": M.Scene}
    for k in fields:
        s, f = k.split(".")
        assert int(out[k]) == getattr(cls[s], f).offset, k


def test_enums_and_version(abi):
    import mdz_amd as M
    out, _ = abi
    assert out["enum.path"].split() == [str(x) for x in (M.PATH_MPFR, M.PATH_LD, M.PATH_DOUBLE,
                                                          M.PATH_GMP)]
    assert out["enum.fractal"].split() == [str(x) for x in (M.MANDELBROT, M.BURNING_SHIP,
                                                            M.GENERALIZED_CELTIC, M.VARIANT)]
    assert int(out["abi"]) == M.lib().mdzg_abi_version()


def test_calls_agree_with_the_ctypes_mirror(abi):
    import mdz_amd as M
    out, png = abi
    rc, *rect = out["rect"].split()
    assert rc == "0" and tuple(rect) == M.rect_from_center("-0.75", "0", "3.0", 128, 240, 180)
    f = os.path.join(MDZ, "floral.mdz")
    d = M.load_scene(f)
    s = out["scene"].split()
    assert s[:6] == ["0", str(d["fractal"]), str(d["path"]), str(d["depth"]),
                     str(d["precision"]), str(d["centre"])] and s[6] == d["c0"]
    rc, w, h = (int(x) for x in out["dims"].split())
    v = M.scene_view(f)
    assert rc == 0 and (w, h) == (v.real_width, v.real_height)
    rc, *srect = out["srect"].split()
    assert rc == "0" and srect == [x.decode() for x in v._keep[:4]]
    p = M.scene_palette(f)
    pr = out["palette"].split()
    assert pr[0] == "0" and int(pr[1]) == p.n and int(pr[2]) == p.offset
    assert float(pr[3]) == p.colour_scale and int(pr[4]) == p.interpolate
    assert (int(pr[5]), int(pr[6])) == (p.colours[0], p.colours[p.n - 1])
    assert out["png"] == "0"
    from tests.test_colour_cpu import png_decode, _rgb_bytes
    w, h, img = png_decode(str(png))
    want = (np.arange(12, dtype=np.uint32) * 0x010203).reshape(3, 4)
    assert (w, h) == (4, 3) and np.array_equal(img, _rgb_bytes(want))
