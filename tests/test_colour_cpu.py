"""Colouring, palettes and PNG without a GPU: the oracle's colouring pinned by
closed forms of MDZ's palette_apply / get_pixel_colour / do_anti_aliasing
(src/palette.c:406-502, src/render.c:108-157) and by an independent float
restatement; palette files against MDZ's gallery maps; the PNG writer by
decoding its files."""
import math
import os
import struct
import zlib

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MDZ = os.path.join(ROOT, "tests", "golden", "mdz")


def _pal(n, seed=0):
    rng = np.random.default_rng(seed)
    return [int(x) for x in rng.integers(0, 1 << 24, n)]


def _ch(c, k):
    return (c >> (8 * k)) & 0xff


# -- an independent restatement on Python floats (IEEE double) -----------------
def _u32(v):
    """(guint32)v as x86-64 converts a non-negative double."""
    return (int(v) & 0xffffffff) if v < 2.0 ** 63 else 0


def _mix(pal, m, v, lo, hi, c):
    wl = c - v
    wh = 1.0 - wl
    a, b = pal[lo % m], pal[hi % m]
    return sum((int(wl * _ch(a, k) + wh * _ch(b, k)) & 0xff) << (8 * k) for k in range(3))


def py_colour(counts, aa, pal, offset, sc, ip):
    n, m = len(pal), len(pal) - 1
    rh, rw = counts.shape
    out = np.zeros((rh // aa, rw // aa), dtype=np.uint32)
    for y in range(rh // aa):
        for x in range(rw // aa):
            if aa == 1:
                raw = int(counts[y, x])
                if raw == 0:
                    continue
                v = raw * sc + offset
                if ip:
                    out[y, x] = _mix(pal, m, v, math.floor(v), math.ceil(v), float(math.ceil(v)))
                else:
                    out[y, x] = pal[_u32(v) % m]
                continue
            s = [0, 0, 0]
            for yi in range(y * aa, y * aa + aa):
                for xi in range(x * aa, x * aa + aa):
                    v = int(counts[yi, xi]) * sc
                    if v != 0.0:
                        v += offset
                    if ip:
                        c = _mix(pal, m, v, _u32(math.floor(v)), _u32(math.ceil(v)),
                                 float(math.ceil(v)))
                    else:
                        c = 0 if v == 0.0 else pal[_u32(v) % m]
                    for k in range(3):
                        s[k] += _ch(c, k)
            out[y, x] = sum((s[k] // (aa * aa)) << (8 * k) for k in range(3))
    return out


def _counts(shape, seed, hi=100000):
    rng = np.random.default_rng(seed)
    c = rng.integers(0, hi, shape).astype(np.uint32)
    c[rng.random(shape) < 0.2] = 0
    return c


# -- closed forms -------------------------------------------------------------
@pytest.mark.parametrize("n,offset", [(2, 0), (3, 7), (256, 148), (17, 255)])
def test_plain_palette_indexes_modulo_n_minus_1(oracle, n, offset):
    pal = _pal(n, n)
    c = _counts((6, 40), n)
    got = oracle.colour(c, 1, pal, offset=offset)
    want = np.array([[pal[(int(r) + offset) % (n - 1)] if r else 0 for r in row] for row in c])
    assert np.array_equal(got, want)


def test_interpolation_at_whole_values_is_the_upper_entry(oracle):
    """colour_scale 1: v is whole, weight 0 on floor(v), 1 on ceil(v)."""
    pal = _pal(256, 1)
    c = _counts((5, 30), 1)
    got = oracle.colour(c, 1, pal, offset=9, interpolate=True)
    want = np.array([[(pal[(int(r) + 9) % 255] & 0xffffff) if r else 0 for r in row] for row in c])
    assert np.array_equal(got, want)


def test_interpolation_halfway_is_the_truncated_mean(oracle):
    """colour_scale 1/2, odd counts: v = k + 1/2, each channel (a + b) >> 1."""
    pal = _pal(100, 2)
    c = (2 * np.arange(1, 61, dtype=np.uint32) + 1).reshape(3, 20)
    got = oracle.colour(c, 1, pal, colour_scale=0.5, interpolate=True)
    m = 99
    for (y, x), r in np.ndenumerate(c):
        k = int(r) // 2
        a, b = pal[k % m], pal[(k + 1) % m]
        assert got[y, x] == sum(((_ch(a, i) + _ch(b, i)) >> 1) << (8 * i) for i in range(3))


def test_index_wraps_like_the_x86_conversion(oracle):
    """(guint32)v of v >= 2^32 keeps the low 32 bits of the 64-bit truncation."""
    pal = _pal(256, 3)
    raw = np.array([[(1 << 20) + 3, (1 << 20) + 5]], dtype=np.uint32)
    got = oracle.colour(raw, 1, pal, colour_scale=float(1 << 20))
    assert list(got[0]) == [pal[((3 << 20) % (1 << 32)) % 255], pal[(5 << 20) % 255]]


@pytest.mark.parametrize("aa", [2, 3])
def test_antialias_of_a_constant_block_is_its_colour(oracle, aa):
    pal = _pal(256, 4)
    base = _counts((4, 7), 4)
    c = np.kron(base, np.ones((aa, aa), dtype=np.uint32))
    for ip in (False, True):
        got = oracle.colour(c, aa, pal, offset=3, interpolate=ip)
        want = np.array([[(pal[(int(r) + 3) % 255] & 0xffffff) if r else
                          ((pal[0] & 0xffffff) if ip else 0) for r in row] for row in base])
        assert np.array_equal(got, want), ip


def test_antialias_averages_and_truncates(oracle):
    """One 2x2 block of four known colours: per channel (sum of 4) // 4."""
    pal = [0x000000, 0x0a0b0c, 0xff0001, 0x030507, 0x102030]
    c = np.array([[1, 2], [3, 0]], dtype=np.uint32)
    got = oracle.colour(c, 2, pal)
    cols = [pal[1], pal[2], pal[3], 0]
    assert got[0, 0] == sum((sum(_ch(x, k) for x in cols) // 4) << (8 * k) for k in range(3))


def test_interpolating_row_path_refuses_mdz_undefined_indexes(oracle):
    with pytest.raises(ValueError):
        oracle.colour(np.array([[2000000]], dtype=np.uint32), 1, _pal(10), colour_scale=2000.0,
                      interpolate=True)


# -- random coverage against the float restatement ------------------------------
@pytest.mark.parametrize("aa", [1, 2, 3])
@pytest.mark.parametrize("ip", [False, True])
@pytest.mark.parametrize("n,offset,sc", [(256, 0, 1.0), (256, 148, 0.02539073500000000108),
                                         (37, 12, 0.18554686500000000526), (2, 1, 3.7),
                                         (200, 255, 1e-3)])
def test_oracle_equals_float_restatement(oracle, aa, ip, n, offset, sc):
    pal = _pal(n, aa + 3 * n)
    c = _counts((6 * aa, 9 * aa), aa * 7 + n)
    assert np.array_equal(oracle.colour(c, aa, pal, offset=offset, colour_scale=sc,
                                        interpolate=ip), py_colour(c, aa, pal, offset, sc, ip))


# -- palette files ---------------------------------------------------------------
def _map_rows(path):
    rows = []
    for line in open(path):
        p = line.split()
        if len(p) < 3:
            break
        rows.append(tuple(int(x) for x in p[:3]))
    return rows[:256]


@pytest.mark.parametrize("name", ["nice.map", "purple_orange.map"])
def test_palette_files(name):
    import mdz_amd as M
    p = M.load_palette(os.path.join(MDZ, name))
    rows = _map_rows(os.path.join(MDZ, name))
    assert p.n == len(rows) == 256
    assert list(M.palette_colours(p)) == [r | (g << 8) | (b << 16) for r, g, b in rows]


def test_palette_read_rules(tmp_path):
    """Reading stops at the first line that is not three integers and after
    256 colours; values are packed with RGB() as given."""
    import mdz_amd as M
    f = tmp_path / "a.map"
    f.write_text(" 1 2 3\n\t4 5 6 extra\n300 -1 0\nnot a colour\n7 8 9\n")
    p = M.load_palette(str(f))
    assert p.n == 3
    assert list(M.palette_colours(p)) == [0x030201, 0x060504, (300 | (-1 << 8)) & 0xffffffff]
    f.write_text("".join(f"{i} 0 0\n" for i in range(300)))
    assert M.load_palette(str(f)).n == 256
    f.write_text("no colours\n")
    with pytest.raises(M.MdzgError):
        M.load_palette(str(f))


def test_scene_palettes():
    import mdz_amd as M
    p = M.scene_palette(os.path.join(MDZ, "floral.mdz"))
    rows = []
    lines = open(os.path.join(MDZ, "floral.mdz")).read().split("\ndata\n", 1)[1].splitlines()
    for line in lines[:256]:
        r, g, b = (int(x) for x in line.split())
        rows.append(r | (g << 8) | (b << 16))
    assert (p.n, p.offset, p.colour_scale, p.interpolate) == (256, 148, 1.0, 0)
    assert list(M.palette_colours(p)) == rows
    h = M.scene_palette(os.path.join(MDZ, "honeytrace.mdz"))
    assert (h.offset, h.interpolate) == (144, 1)
    # spiralblues names a palette file of a gnofract4d install: not here
    assert M.load_scene(os.path.join(MDZ, "spiralblues.mdz"))["palette_n"] == 0
    with pytest.raises(M.MdzgError):
        M.scene_palette(os.path.join(MDZ, "spiralblues.mdz"))


# -- PNG ---------------------------------------------------------------------------
def png_decode(path):
    """(width, height, [h, w, 3] uint8) of an 8-bit RGB non-interlaced PNG,
    checking every chunk CRC."""
    b = open(path, "rb").read()
    assert b[:8] == b"\x89PNG\r\n\x1a\n"
    i, idat, ihdr, types = 8, b"", None, []
    while i < len(b):
        ln, = struct.unpack(">I", b[i:i + 4])
        t, d = b[i + 4:i + 8], b[i + 8:i + 8 + ln]
        crc, = struct.unpack(">I", b[i + 8 + ln:i + 12 + ln])
        assert crc == zlib.crc32(t + d) & 0xffffffff
        types.append(t)
        if t == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", d)
        elif t == b"IDAT":
            idat += d
        i += 12 + ln
    assert types[0] == b"IHDR" and types[-1] == b"IEND"
    w, h, depth, ctype, comp, filt, inter = ihdr
    assert (depth, ctype, comp, filt, inter) == (8, 2, 0, 0, 0)
    raw = zlib.decompress(idat)
    stride = 3 * w
    assert len(raw) == h * (stride + 1)
    img = np.zeros((h, stride), dtype=np.int64)
    prev = np.zeros(stride, dtype=np.int64)
    for y in range(h):
        ft = raw[y * (stride + 1)]
        cur = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        out = np.zeros(stride, dtype=np.int64)
        for x in range(stride):
            a = out[x - 3] if x >= 3 else 0
            bb = prev[x]
            c = prev[x - 3] if x >= 3 else 0
            if ft == 0:
                pr = 0
            elif ft == 1:
                pr = a
            elif ft == 2:
                pr = bb
            elif ft == 3:
                pr = (a + bb) // 2
            else:
                p = a + bb - c
                pa, pb, pc = abs(p - a), abs(p - bb), abs(p - c)
                pr = a if pa <= pb and pa <= pc else (bb if pb <= pc else c)
            out[x] = (cur[x] + pr) & 0xff
        img[y] = out
        prev = out
    return w, h, img.reshape(h, w, 3).astype(np.uint8)


def _rgb_bytes(rgb):
    return np.stack([(rgb >> (8 * k)) & 0xff for k in range(3)], axis=-1).astype(np.uint8)


@pytest.mark.parametrize("w,h", [(1, 1), (7, 5), (64, 48), (301, 3)])
def test_png_round_trip(tmp_path, w, h):
    import mdz_amd as M
    rng = np.random.default_rng(w * h)
    rgb = rng.integers(0, 1 << 32, (h, w), dtype=np.uint64).astype(np.uint32)
    rgb[: h // 2, : w // 2] = 0x123456                 # flat area: Sub/Up/Paeth rows
    rgb[h // 2:, :] = (np.arange(w, dtype=np.uint32) * 0x010203)[None, :]
    f = tmp_path / "a.png"
    M.png_write(str(f), rgb)
    W, H, img = png_decode(str(f))
    assert (W, H) == (w, h)
    assert np.array_equal(img, _rgb_bytes(rgb))     # the byte above blue is dropped


def test_png_bad_path():
    import mdz_amd as M
    with pytest.raises(M.MdzgError):
        M.png_write("/nonexistent-dir/a.png", np.zeros((2, 2), dtype=np.uint32))


def test_colour_helpers_on_the_host(tmp_path):
    """The colour kernel's remainder and conversion helpers, built for the
    host from the same header, against % and the truncation they restate."""
    import subprocess
    exe = tmp_path / "colour_selftest"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-o", str(exe),
                    os.path.join(ROOT, "tests", "colour_selftest.cpp")], check=True)
    r = subprocess.run([str(exe), "100000"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
