"""ctypes binding of the oracle (TEST INFRASTRUCTURE ONLY).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, as the checker.  The product (mdz_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libfrac_ref.so")

MANDELBROT, BURNING_SHIP, GENERALIZED_CELTIC, VARIANT = 0, 1, 2, 3
FAMILY_MANDEL, FAMILY_JULIA = 0, 1
PATH_MPFR, PATH_LD, PATH_DOUBLE, PATH_GMP = 0, 1, 2, 3


class Frame(ctypes.Structure):
    _fields_ = [
        ("fractal", ctypes.c_int), ("family", ctypes.c_int), ("path", ctypes.c_int),
        ("real_width", ctypes.c_int), ("real_height", ctypes.c_int),
        ("precision", ctypes.c_long), ("depth", ctypes.c_longlong),
        ("xmin", ctypes.c_char_p), ("xmax", ctypes.c_char_p),
        ("ymax", ctypes.c_char_p), ("width", ctypes.c_char_p),
        ("julia_re", ctypes.c_char_p), ("julia_im", ctypes.c_char_p),
        ("line0", ctypes.c_int), ("nlines", ctypes.c_int), ("line_step", ctypes.c_int),
        ("col0", ctypes.c_int), ("ncols", ctypes.c_int),
    ]


_lib = None


def build():
    """Compile liboracle.so (and oracle/_ref when the reference tree exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-B", "-C", HERE, "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle/liboracle.so is not built (run make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_render.argtypes = [ctypes.POINTER(Frame), ctypes.c_int,
                                 ctypes.c_void_p, ctypes.POINTER(ctypes.c_double)]
        L.orc_render.restype = ctypes.c_int
        L.orc_use_reference.argtypes = [ctypes.c_char_p]
        L.orc_use_reference.restype = ctypes.c_int
        L.orc_rect_from_center.argtypes = [ctypes.c_char_p] * 3 + [
            ctypes.c_long, ctypes.c_int, ctypes.c_int] + [ctypes.c_char_p] * 4 + [ctypes.c_int]
        L.orc_rect_from_center.restype = ctypes.c_int
        L.orc_rect_from_center_coords.argtypes = L.orc_rect_from_center.argtypes
        L.orc_rect_from_center_coords.restype = ctypes.c_int
        L.orc_formula_mpfr.argtypes = [ctypes.c_int, ctypes.c_longlong, ctypes.c_long] + \
            [ctypes.c_char_p] * 6 + [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_formula_mpfr.restype = ctypes.c_longlong
        L.orc_formula_ld.argtypes = [ctypes.c_int, ctypes.c_longlong] + [ctypes.c_char_p] * 6
        L.orc_formula_ld.restype = ctypes.c_longlong
        L.orc_formula_d.argtypes = [ctypes.c_int, ctypes.c_longlong] + [ctypes.c_double] * 6
        L.orc_formula_d.restype = ctypes.c_longlong
        L.orc_pixel_c.argtypes = [ctypes.POINTER(Frame), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_pixel_c.restype = ctypes.c_int
        L.orc_rn_op.argtypes = [ctypes.c_int, ctypes.c_long, ctypes.c_char_p,
                                ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int]
        L.orc_rn_op.restype = ctypes.c_int
        L.orc_colour.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_void_p]
        L.orc_colour.restype = ctypes.c_int
        _lib = L
    return _lib


def use_reference(on=True, build="O0"):
    """Switch the ld/MPFR/GMP loops to the compiled reference loops (oracle/_ref):
    build "O0" as the reference builds them (-ggdb), "O2" optimised."""
    if on:
        path = REF_PATH if build == "O0" else REF_PATH.replace(".so", "_O2.so")
        if not os.path.exists(path):
            raise RuntimeError(f"{path} not built")
        rc = lib().orc_use_reference(path.encode())
    else:
        rc = lib().orc_use_reference(None)
    if rc != 0:
        raise RuntimeError("could not load the reference loops")


def _b(s):
    return None if s is None else (s.encode() if isinstance(s, str) else s)


def rect_from_center(cx, cy, size, precision, user_w, user_h, coords=False):
    """(xmin, xmax, ymax, width) hex strings at max(p, 80) bits (as image_info
    holds them), or with coords=True at the coords precision (as coords holds
    them: the input of the GMP path)."""
    cap = 4096
    bufs = [ctypes.create_string_buffer(cap) for _ in range(4)]
    fn = lib().orc_rect_from_center_coords if coords else lib().orc_rect_from_center
    rc = fn(_b(cx), _b(cy), _b(size), precision, user_w, user_h, *bufs, cap)
    if rc != 0:
        raise ValueError("bad rect arguments")
    return tuple(b.value.decode() for b in bufs)


def render(rect, *, precision, depth, width, height, fractal=MANDELBROT,
           family=FAMILY_MANDEL, path=PATH_MPFR, julia=(None, None),
           lines=None, cols=None, threads=1, timed=False):
    """Render rows/columns of a frame.  rect = (xmin, xmax, ymax, width) strings.

    lines: (line0, nlines, line_step) or None for all rows;
    cols:  (col0, ncols) or None for all columns.
    Returns an int32 array [nlines, ncols] (and the elapsed seconds if timed).
    """
    line0, nlines, step = lines if lines is not None else (0, height, 1)
    col0, ncols = cols if cols is not None else (0, width)
    keep = [_b(s) for s in rect] + [_b(julia[0]), _b(julia[1])]
    fr = Frame(fractal, family, path, width, height, precision, depth,
               keep[0], keep[1], keep[2], keep[3], keep[4], keep[5],
               line0, nlines, step, col0, ncols)
    out = np.zeros((nlines, ncols), dtype=np.int32)
    el = ctypes.c_double(0.0)
    rc = lib().orc_render(ctypes.byref(fr), threads, out.ctypes.data, ctypes.byref(el))
    if rc != 0:
        raise ValueError("oracle rejected the frame")
    return (out, el.value) if timed else out


def formula_mpfr(fractal, depth, precision, wim, wre, c_im, c_re, wim2, wre2):
    cap = 4096
    a = ctypes.create_string_buffer(cap)
    b = ctypes.create_string_buffer(cap)
    n = lib().orc_formula_mpfr(fractal, depth, precision, *[_b(v) for v in
                               (wim, wre, c_im, c_re, wim2, wre2)], a, b, cap)
    return n, a.value.decode(), b.value.decode()


def rn_op(op, precision, a, b):
    """MPFR RNDN add (0), sub (1) or mul (2) at `precision`; hex result."""
    buf = ctypes.create_string_buffer(8192)
    if lib().orc_rn_op(op, precision, _b(a), _b(b), buf, 8192) != 0:
        raise ValueError("bad operands")
    return buf.value.decode()


def formula_ld(fractal, depth, wim, wre, c_im, c_re, wim2, wre2):
    """The x87 long double loop on explicit start values, given as strings
    (hex-float or decimal) that the C side converts exactly with strtold."""
    return lib().orc_formula_ld(fractal, depth, *[_b(v) for v in
                                                  (wim, wre, c_im, c_re, wim2, wre2)])


def formula_d(fractal, depth, wim, wre, c_im, c_re, wim2, wre2):
    return lib().orc_formula_d(fractal, depth, wim, wre, c_im, c_re, wim2, wre2)


def pixel_c(rect, *, precision, width, height, line, ix, path=PATH_MPFR):
    """(x, y) of one pixel as the line worker computes it (strings)."""
    keep = [_b(s) for s in rect]
    fr = Frame(0, 0, path, width, height, precision, 1, keep[0], keep[1], keep[2], keep[3],
               None, None, 0, 0, 1, 0, 0)
    a = ctypes.create_string_buffer(4096)
    b = ctypes.create_string_buffer(4096)
    if lib().orc_pixel_c(ctypes.byref(fr), line, ix, a, b, 4096) != 0:
        raise ValueError("bad pixel")
    return a.value.decode(), b.value.decode()


def colour(counts, aa, colours, *, offset=0, colour_scale=1.0, interpolate=False):
    """rgb_data [h, w] (uint32, r | g << 8 | b << 16) of a [h*aa, w*aa] count
    array, as MDZ's palette_apply / do_anti_aliasing compute it."""
    c = np.ascontiguousarray(counts, dtype=np.uint32)
    rh, rw = c.shape
    pal = np.zeros(256, dtype=np.uint32)
    cols = np.asarray(colours, dtype=np.uint32)
    pal[:len(cols)] = cols
    out = np.zeros((rh // aa, rw // aa), dtype=np.uint32)
    rc = lib().orc_colour(c.ctypes.data, rw // aa, rh // aa, aa, pal.ctypes.data, len(cols),
                          offset, colour_scale, int(interpolate), out.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle rejected the colouring ({rc})")
    return out
