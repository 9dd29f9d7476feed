"""mdz_amd -- MI355X (gfx950) escape-time engine for MDZ, Python binding.

A thin ctypes layer over the C ABI of include/mdzg.h (libmdzg.so, built in
tree by mdz_amd.build).  Loading fails loudly when the library is missing:
there is no CPU path behind these calls.

Names follow the reference: a *view* is the set of image_info fields the line
workers read (src/image_info.h:63-122), a *tile* a rectangle of rows and
columns of a view, counts are the per-pixel iteration counts the line
workers store in raw_data (0 = never escaped).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdzg.so")

MANDELBROT, BURNING_SHIP, GENERALIZED_CELTIC, VARIANT = 0, 1, 2, 3
FAMILY_MANDEL, FAMILY_JULIA = 0, 1
PATH_MPFR, PATH_LD, PATH_DOUBLE, PATH_GMP = 0, 1, 2, 3
BATCH_PERSISTENT, BATCH_PER_TILE = 0, 1
SERVE_BATCHED, SERVE_PER_REQUEST = 0, 1
MAX_PRECISION = 1024

ERRORS = {0: "ok", -1: "invalid argument", -2: "out of memory", -3: "HIP device error",
          -4: "render stopped", -5: "admission refused", -6: "coordinate string did not parse",
          -7: "precision or size out of range", -8: "call out of sequence"}


class MdzgError(RuntimeError):
    def __init__(self, code, what=""):
        super().__init__(f"{what}: {ERRORS.get(code, code)} ({code})")
        self.code = code


class View(ctypes.Structure):
    _fields_ = [
        ("fractal", ctypes.c_int), ("family", ctypes.c_int), ("path", ctypes.c_int),
        ("real_width", ctypes.c_int), ("real_height", ctypes.c_int),
        ("precision", ctypes.c_long), ("depth", ctypes.c_longlong),
        ("xmin", ctypes.c_char_p), ("xmax", ctypes.c_char_p),
        ("ymax", ctypes.c_char_p), ("width", ctypes.c_char_p),
        ("julia_re", ctypes.c_char_p), ("julia_im", ctypes.c_char_p),
    ]


class Tile(ctypes.Structure):
    _fields_ = [("view", View), ("line0", ctypes.c_int), ("nlines", ctypes.c_int),
                ("col0", ctypes.c_int), ("ncols", ctypes.c_int),
                ("counts", ctypes.POINTER(ctypes.c_uint32))]


class Scene(ctypes.Structure):
    _fields_ = [
        ("version_major", ctypes.c_int), ("version_minor", ctypes.c_int),
        ("version_revision", ctypes.c_int), ("family", ctypes.c_int), ("fractal", ctypes.c_int),
        ("depth", ctypes.c_longlong), ("aspect", ctypes.c_double),
        ("colour_scale", ctypes.c_double), ("colour_interpolate", ctypes.c_int),
        ("multi_precision", ctypes.c_int), ("rounding", ctypes.c_int), ("path", ctypes.c_int),
        ("precision", ctypes.c_long), ("centre", ctypes.c_int),
        ("c0", ctypes.c_char_p), ("c1", ctypes.c_char_p), ("c2", ctypes.c_char_p),
        ("julia_re", ctypes.c_char_p), ("julia_im", ctypes.c_char_p),
        ("palette_offset", ctypes.c_int),
        ("palette_n", ctypes.c_int), ("palette", ctypes.c_uint32 * 256),
    ]


class Palette(ctypes.Structure):
    _fields_ = [("colours", ctypes.c_uint32 * 256), ("n", ctypes.c_int),
                ("offset", ctypes.c_int), ("colour_scale", ctypes.c_double),
                ("interpolate", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("wall_ms", ctypes.c_double),
                ("pixel_iters", ctypes.c_double), ("launches", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TuneResult(ctypes.Structure):
    _fields_ = [("kernel_id", ctypes.c_int), ("blocks_per_cu", ctypes.c_int),
                ("chunk", ctypes.c_int), ("best_ms", ctypes.c_double),
                ("default_ms", ctypes.c_double), ("candidates", ctypes.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


OPT_GRAPHS, OPT_BAND_PIXELS, OPT_BAND_INFLIGHT = 1, 2, 3
# operations of Context.op_probe (mdzg_op_probe)
OPP_LOOP_ONLY, OPP_PRODUCT, OPP_SQUARE, OPP_ADD_SUB, OPP_SQUARES_DIFF, OPP_ESCAPE, \
    OPP_ITERATION = range(7)


class ServerCfg(ctypes.Structure):
    _fields_ = [("mode", ctypes.c_int), ("hbm_budget", ctypes.c_size_t),
                ("occupancy_budget", ctypes.c_double), ("slice_iters", ctypes.c_int),
                ("max_requests", ctypes.c_int)]


class ReqInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("submit_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("done_ms", ctypes.c_double),
                ("latency_ms", ctypes.c_double), ("pixel_iters", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ServeStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("admitted", ctypes.c_uint64),
                ("completed", ctypes.c_uint64), ("slices", ctypes.c_uint64),
                ("active", ctypes.c_uint64), ("pending", ctypes.c_uint64),
                ("pixel_iters", ctypes.c_double), ("hbm_in_use", ctypes.c_size_t),
                ("hbm_peak", ctypes.c_size_t)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_lib = None


def lib():
    """The loaded libmdzg.so (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m mdz_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.POINTER
        vp = ctypes.c_void_p
        L.mdzg_abi_version.restype = ctypes.c_int
        L.mdzg_strerror.argtypes = [ctypes.c_int]
        L.mdzg_strerror.restype = ctypes.c_char_p
        L.mdzg_create.argtypes = [ctypes.c_int, P(vp)]
        L.mdzg_destroy.argtypes = [vp]
        L.mdzg_destroy.restype = None
        L.mdzg_rect_from_center.argtypes = [ctypes.c_char_p] * 3 + [
            ctypes.c_long, ctypes.c_int, ctypes.c_int] + [ctypes.c_char_p] * 4 + [ctypes.c_int]
        L.mdzg_rect_from_center_coords.argtypes = L.mdzg_rect_from_center.argtypes
        L.mdzg_scene_load.argtypes = [ctypes.c_char_p, P(Scene)]
        L.mdzg_scene_free.argtypes = [P(Scene)]
        L.mdzg_scene_free.restype = None
        L.mdzg_scene_dims.argtypes = [P(Scene), P(ctypes.c_int), P(ctypes.c_int)]
        L.mdzg_scene_palette.argtypes = [P(Scene), P(Palette)]
        L.mdzg_palette_load.argtypes = [ctypes.c_char_p, P(Palette)]
        L.mdzg_colour.argtypes = [vp, P(ctypes.c_uint32), ctypes.c_int, ctypes.c_int,
                                  ctypes.c_int, P(Palette), P(ctypes.c_uint32), P(Stats)]
        L.mdzg_render_image.argtypes = [vp, P(View), ctypes.c_int, P(Palette),
                                        P(ctypes.c_uint32), P(ctypes.c_uint32), P(Stats)]
        L.mdzg_png_write.argtypes = [ctypes.c_char_p, P(ctypes.c_uint32), ctypes.c_int,
                                     ctypes.c_int]
        L.mdzg_server_submit_image.argtypes = [vp, P(Tile), ctypes.c_int, P(Palette),
                                               P(ctypes.c_uint32), P(ctypes.c_uint64)]
        L.mdzg_scene_rect.argtypes = [P(Scene), ctypes.c_int, ctypes.c_int, ctypes.c_int] + \
            [ctypes.c_char_p] * 4 + [ctypes.c_int]
        L.mdzg_render_tile.argtypes = [vp, P(Tile), P(Stats)]
        L.mdzg_render_tile_noalloc.argtypes = [vp, P(Tile), P(Stats)]
        L.mdzg_render_batch.argtypes = [vp, P(Tile), ctypes.c_int, ctypes.c_int, P(Stats)]
        L.mdzg_submit.argtypes = [vp, P(Tile), ctypes.c_int, P(vp)]
        L.mdzg_poll.argtypes = [vp]
        L.mdzg_wait.argtypes = [vp, P(Stats)]
        L.mdzg_cancel.argtypes = [vp]
        L.mdzg_job_free.argtypes = [vp]
        L.mdzg_job_free.restype = None
        L.mdzg_frame_open.argtypes = [vp, P(View), P(ctypes.c_int32), P(vp)]
        L.mdzg_frame_calculate_line.argtypes = [vp, ctypes.c_int]
        L.mdzg_frame_stop.argtypes = [vp]
        L.mdzg_frame_stop.restype = None
        L.mdzg_frame_close.argtypes = [vp, P(Stats)]
        L.mdzg_server_create.argtypes = [vp, P(ServerCfg), P(vp)]
        L.mdzg_server_submit.argtypes = [vp, P(Tile), P(ctypes.c_uint64)]
        L.mdzg_server_poll.argtypes = [vp, ctypes.c_uint64]
        L.mdzg_server_wait.argtypes = [vp, ctypes.c_uint64, P(ReqInfo)]
        L.mdzg_server_info.argtypes = [vp, ctypes.c_uint64, P(ReqInfo)]
        L.mdzg_server_stats.argtypes = [vp, P(ServeStats)]
        L.mdzg_server_destroy.argtypes = [vp]
        L.mdzg_server_destroy.restype = None
        L.mdzg_option_set.argtypes = [vp, ctypes.c_int, ctypes.c_longlong]
        L.mdzg_option_get.argtypes = [vp, ctypes.c_int, P(ctypes.c_longlong)]
        L.mdzg_autotune.argtypes = [vp, P(Tile), ctypes.c_int, ctypes.c_int, P(TuneResult)]
        L.mdzg_tune_save.argtypes = [vp, ctypes.c_char_p]
        L.mdzg_tune_load.argtypes = [vp, ctypes.c_char_p]
        L.mdzg_rate_probe.argtypes = [vp, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
        L.mdzg_op_probe.argtypes = [vp, P(Tile), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                    P(ctypes.c_double)]
        L.mdzg_kernel_id.argtypes = [ctypes.c_int, ctypes.c_long]
        L.mdzg_tune_set.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.mdzg_tune_get.argtypes = [vp, ctypes.c_int] + [P(ctypes.c_int)] * 3
        _lib = L
    return _lib


def _b(s):
    if s is None:
        return None
    return s.encode() if isinstance(s, str) else s


def _check(rc, what):
    if rc < 0:
        raise MdzgError(rc, what)
    return rc


def rect_from_center(cx, cy, size, precision, user_w, user_h, coords=False):
    """img->xmin, xmax, ymax, width (hex strings at max(p,80) bits) of a centre view,
    as the command-line path computes them (src/coords.c:204-313); with
    coords=True the same rect at the coords precision, the GMP path's input
    (src/coords.c:317-324)."""
    cap = 4096
    bufs = [ctypes.create_string_buffer(cap) for _ in range(4)]
    fn = lib().mdzg_rect_from_center_coords if coords else lib().mdzg_rect_from_center
    _check(fn(_b(cx), _b(cy), _b(size), precision, user_w, user_h, *bufs, cap),
           "rect_from_center")
    return tuple(b.value.decode() for b in bufs)


def kernel_id(path, precision):
    """Id of the kernel a view on (path, precision) runs on."""
    return _check(lib().mdzg_kernel_id(path, precision), "kernel_id")


def make_palette(colours, *, offset=0, colour_scale=1.0, interpolate=False):
    """A Palette from a sequence of RGB(r, g, b) colours (r | g << 8 | b << 16)
    or of (r, g, b) triples."""
    p = Palette()
    cols = [c if isinstance(c, int) else (c[0] | (c[1] << 8) | (c[2] << 16)) for c in colours]
    if not 2 <= len(cols) <= 256:
        raise ValueError("a palette has 2..256 colours")
    for i, c in enumerate(cols):
        p.colours[i] = c & 0xffffffff
    p.n, p.offset, p.colour_scale, p.interpolate = len(cols), offset, colour_scale, int(interpolate)
    return p


def load_palette(path, *, offset=0, colour_scale=1.0, interpolate=False):
    """A .map palette file (mdzg_palette_load)."""
    p = make_palette([0, 0], offset=offset, colour_scale=colour_scale, interpolate=interpolate)
    _check(lib().mdzg_palette_load(_b(path), ctypes.byref(p)), "palette_load")
    return p


def scene_palette(path):
    """The palette of a .mdz file with its offset, colour scale and interpolation."""
    sc = Scene()
    _check(lib().mdzg_scene_load(_b(path), ctypes.byref(sc)), "scene_load")
    try:
        p = Palette()
        _check(lib().mdzg_scene_palette(ctypes.byref(sc), ctypes.byref(p)), "scene_palette")
    finally:
        lib().mdzg_scene_free(ctypes.byref(sc))
    return p


def palette_colours(p):
    return np.ctypeslib.as_array(p.colours)[:p.n].copy()


def png_write(path, rgb):
    """Write an [h, w] uint32 rgb_data array as an 8-bit RGB PNG (mdzg_png_write)."""
    a = np.ascontiguousarray(rgb, dtype=np.uint32)
    h, w = a.shape
    _check(lib().mdzg_png_write(_b(path), a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                w, h), "png_write")


def load_scene(path):
    """The settings of a .mdz file (mdzg_scene_load) as a dict."""
    sc = Scene()
    _check(lib().mdzg_scene_load(_b(path), ctypes.byref(sc)), "scene_load")
    try:
        d = {k: getattr(sc, k) for k, _ in Scene._fields_}
        d["palette"] = np.ctypeslib.as_array(sc.palette)[:sc.palette_n].copy()
        for k in ("c0", "c1", "c2", "julia_re", "julia_im"):
            d[k] = d[k].decode() if d[k] is not None else None
    finally:
        lib().mdzg_scene_free(ctypes.byref(sc))
    return d


SCENE_FIT_RECT = 1


def scene_view(path, width=0, height=0, aa=1, fit_rect=False):
    """View of a .mdz file as MDZ's command line renders it: width x height
    (0 = as main.c's init_misc fills it in, mdzg_scene_dims) with aa x aa
    samples per pixel, on the rect of mdzg_scene_rect (fit_rect: a
    corner-form file's own rect refitted, MDZG_SCENE_FIT_RECT)."""
    sc = Scene()
    _check(lib().mdzg_scene_load(_b(path), ctypes.byref(sc)), "scene_load")
    try:
        w, h = ctypes.c_int(width), ctypes.c_int(height)
        _check(lib().mdzg_scene_dims(ctypes.byref(sc), ctypes.byref(w), ctypes.byref(h)),
               "scene_dims")
        cap = 4096
        bufs = [ctypes.create_string_buffer(cap) for _ in range(4)]
        _check(lib().mdzg_scene_rect(ctypes.byref(sc), w.value, h.value,
                                     SCENE_FIT_RECT if fit_rect else 0, *bufs, cap), "scene_rect")
        rect = tuple(b.value.decode() for b in bufs)
        julia = (sc.julia_re.decode(), sc.julia_im.decode()) if sc.julia_re else (None, None)
        v = make_view(rect, precision=sc.precision, depth=sc.depth, width=w.value * aa,
                      height=h.value * aa, fractal=sc.fractal, family=sc.family, path=sc.path,
                      julia=julia)
    finally:
        lib().mdzg_scene_free(ctypes.byref(sc))
    return v


def make_view(rect, *, precision, depth, width, height, fractal=MANDELBROT,
              family=FAMILY_MANDEL, path=PATH_MPFR, julia=(None, None)):
    """View for rect = (xmin, xmax, ymax, width) strings.  The returned View keeps
    its strings alive through the `_keep` attribute."""
    keep = [_b(s) for s in rect] + [_b(julia[0]), _b(julia[1])]
    v = View(fractal, family, path, width, height, precision, depth, *keep)
    v._keep = keep
    return v


class Context:
    """One device context (mdzg_create)."""

    def __init__(self, device=0):
        self._c = ctypes.c_void_p()
        _check(lib().mdzg_create(device, ctypes.byref(self._c)), "mdzg_create")

    def close(self):
        if self._c:
            lib().mdzg_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @staticmethod
    def _tile(view, lines, cols, fill=0):
        line0, nlines = lines if lines is not None else (0, view.real_height)
        col0, ncols = cols if cols is not None else (0, view.real_width)
        out = np.full((nlines, ncols), fill, dtype=np.uint32)
        t = Tile(view, line0, nlines, col0, ncols,
                 out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)))
        return t, out

    def render(self, view, lines=None, cols=None):
        """Counts of a rectangle of a view: uint32 [nlines, ncols]; plus stats."""
        t, out = self._tile(view, lines, cols)
        st = Stats()
        _check(lib().mdzg_render_tile(self._c, ctypes.byref(t), ctypes.byref(st)), "render")
        return out, st.as_dict()

    def render_noalloc(self, view, counts_ptr, lines=None, cols=None):
        """mdzg_render_tile_noalloc: the counts of a rectangle written to device
        memory at address counts_ptr (nlines * ncols uint32 on this context's
        device), with no device allocation; returns the stats."""
        line0, nlines = lines if lines is not None else (0, view.real_height)
        col0, ncols = cols if cols is not None else (0, view.real_width)
        t = Tile(view, line0, nlines, col0, ncols,
                 ctypes.cast(ctypes.c_void_p(counts_ptr), ctypes.POINTER(ctypes.c_uint32)))
        st = Stats()
        _check(lib().mdzg_render_tile_noalloc(self._c, ctypes.byref(t), ctypes.byref(st)),
               "render_noalloc")
        return st.as_dict()

    def colour(self, counts, aa, palette, stats=False):
        """rgb_data [h, w] uint32 of a [h*aa, w*aa] count array (mdzg_colour);
        with stats=True also the stats of the colouring kernel."""
        c = np.ascontiguousarray(counts, dtype=np.uint32)
        rh, rw = c.shape
        if rw % aa or rh % aa:
            raise ValueError("count array is not a whole number of aa x aa blocks")
        out = np.empty((rh // aa, rw // aa), dtype=np.uint32)
        P = ctypes.POINTER(ctypes.c_uint32)
        st = Stats()
        _check(lib().mdzg_colour(self._c, c.ctypes.data_as(P), rw // aa, rh // aa, aa,
                                 ctypes.byref(palette), out.ctypes.data_as(P),
                                 ctypes.byref(st)), "colour")
        return (out, st.as_dict()) if stats else out

    def render_image(self, view, aa, palette, counts=False):
        """Render a whole view and colour it on the device (mdzg_render_image):
        rgb_data [real_height/aa, real_width/aa] (and the counts when asked), stats."""
        P = ctypes.POINTER(ctypes.c_uint32)
        rgb = np.empty((view.real_height // aa, view.real_width // aa), dtype=np.uint32)
        cnt = np.empty((view.real_height, view.real_width), dtype=np.uint32) if counts else None
        st = Stats()
        _check(lib().mdzg_render_image(self._c, ctypes.byref(view), aa, ctypes.byref(palette),
                                       rgb.ctypes.data_as(P),
                                       cnt.ctypes.data_as(P) if counts else None,
                                       ctypes.byref(st)), "render_image")
        return (rgb, cnt, st.as_dict()) if counts else (rgb, st.as_dict())

    def render_batch(self, specs, mode=BATCH_PERSISTENT):
        """specs: list of (view, lines, cols).  Returns ([counts...], stats)."""
        n = len(specs)
        arr = (Tile * n)()
        outs = []
        for i, (v, lines, cols) in enumerate(specs):
            t, out = self._tile(v, lines, cols)
            arr[i] = t
            outs.append(out)
        st = Stats()
        _check(lib().mdzg_render_batch(self._c, arr, n, mode, ctypes.byref(st)), "render_batch")
        return outs, st.as_dict()

    def tune_get(self, kernel_id):
        """(blocks_per_cu, chunk, max_blocks_per_cu) of a kernel."""
        vals = [ctypes.c_int() for _ in range(3)]
        _check(lib().mdzg_tune_get(self._c, kernel_id, *[ctypes.byref(v) for v in vals]),
               "tune_get")
        return tuple(v.value for v in vals)

    def tune_set(self, kernel_id, blocks_per_cu=0, chunk=0):
        _check(lib().mdzg_tune_set(self._c, kernel_id, blocks_per_cu, chunk), "tune_set")

    def option_set(self, option, value):
        _check(lib().mdzg_option_set(self._c, option, int(value)), "option_set")

    def option_get(self, option):
        v = ctypes.c_longlong()
        _check(lib().mdzg_option_get(self._c, option, ctypes.byref(v)), "option_get")
        return v.value

    def autotune(self, views, reps=2):
        """mdzg_autotune on whole-view probes (a view or a list of views of one
        kernel class): the kept pair and times."""
        views = views if isinstance(views, (list, tuple)) else [views]
        arr = (Tile * len(views))()
        keep = []
        for i, v in enumerate(views):
            t, out = self._tile(v, None, None)
            arr[i] = t
            keep.append(out)
        r = TuneResult()
        _check(lib().mdzg_autotune(self._c, arr, len(views), reps, ctypes.byref(r)), "autotune")
        return r.as_dict()

    def rate_probe(self, kind, blocks_per_cu=8):
        """Lane operations per second of v_mad_u64_u32 (kind 0) or v_fma_f64 (1)."""
        v = ctypes.c_double()
        _check(lib().mdzg_rate_probe(self._c, kind, blocks_per_cu, ctypes.byref(v)),
               "rate_probe")
        return v.value

    def op_probe(self, view, op, warm=1000, reps=1000, lines=None, cols=None):
        """mdzg_op_probe: kernel ms of one RN operation (OPP_*) repeated reps
        times per lane after warm iterations of each pixel's orbit."""
        t, _ = self._tile(view, lines, cols)
        v = ctypes.c_double()
        _check(lib().mdzg_op_probe(self._c, ctypes.byref(t), op, warm, reps, ctypes.byref(v)),
               "op_probe")
        return v.value

    def tune_save(self, path):
        _check(lib().mdzg_tune_save(self._c, _b(path)), "tune_save")

    def tune_load(self, path):
        _check(lib().mdzg_tune_load(self._c, _b(path)), "tune_load")

    def submit(self, specs, fill=0):
        """Asynchronous batch: returns a Job.  The counts buffers start at
        `fill` (a cancelled job leaves unfinished pixels at that value)."""
        return Job(self, specs, fill)


class Job:
    def __init__(self, ctx, specs, fill=0):
        n = len(specs)
        self._arr = (Tile * n)()
        self.counts = []
        for i, (v, lines, cols) in enumerate(specs):
            t, out = Context._tile(v, lines, cols, fill)
            self._arr[i] = t
            self.counts.append(out)
        self._views = [s[0] for s in specs]
        self._j = ctypes.c_void_p()
        _check(lib().mdzg_submit(ctx._c, self._arr, n, ctypes.byref(self._j)), "submit")

    def poll(self):
        return _check(lib().mdzg_poll(self._j), "poll") == 1

    def cancel(self):
        _check(lib().mdzg_cancel(self._j), "cancel")

    def wait(self):
        st = Stats()
        rc = lib().mdzg_wait(self._j, ctypes.byref(st))
        return rc, st.as_dict()

    def free(self):
        if self._j:
            lib().mdzg_job_free(self._j)
            self._j = ctypes.c_void_p()


class Frame:
    """The line-callback adapter: calculate_line(line) fills row `line` of raw
    (int32, real_height x real_width) and returns 1, or 0 once stopped."""

    def __init__(self, ctx, view, raw):
        assert raw.dtype == np.int32 and raw.flags.c_contiguous
        self.view = view
        self.raw = raw
        self._f = ctypes.c_void_p()
        _check(lib().mdzg_frame_open(ctx._c, ctypes.byref(view),
                                     raw.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                     ctypes.byref(self._f)), "frame_open")

    def calculate_line(self, line):
        return lib().mdzg_frame_calculate_line(self._f, line)

    def stop(self):
        lib().mdzg_frame_stop(self._f)

    def close(self):
        st = Stats()
        rc = lib().mdzg_frame_close(self._f, ctypes.byref(st))
        self._f = ctypes.c_void_p()
        return rc, st.as_dict()


class Server:
    """Continuous-batching tile server (mdzg_server_*).  submit() returns an id;
    wait(id) returns (counts, info).  The server keeps each request's tile, view
    and counts buffer alive until it is waited for."""

    def __init__(self, ctx, mode=SERVE_BATCHED, hbm_budget=0, occupancy_budget=0.0,
                 slice_iters=0, max_requests=0):
        self.ctx = ctx
        cfg = ServerCfg(mode, hbm_budget, occupancy_budget, slice_iters, max_requests)
        self._s = ctypes.c_void_p()
        _check(lib().mdzg_server_create(ctx._c, ctypes.byref(cfg), ctypes.byref(self._s)),
               "server_create")
        self._live = {}

    def submit(self, view, lines=None, cols=None):
        t, out = Context._tile(view, lines, cols)
        rid = ctypes.c_uint64()
        _check(lib().mdzg_server_submit(self._s, ctypes.byref(t), ctypes.byref(rid)), "submit")
        self._live[rid.value] = (t, out, view)
        return rid.value

    def submit_image(self, view, aa, palette, lines=None, cols=None, counts=False):
        """An image request (mdzg_server_submit_image): wait(id) returns
        (rgb, info), or ((rgb, counts), info) with counts=True."""
        t, out = Context._tile(view, lines, cols)
        if t.nlines % aa or t.ncols % aa:
            raise ValueError("tile is not a whole number of aa x aa blocks")
        rgb = np.zeros((t.nlines // aa, t.ncols // aa), dtype=np.uint32)
        if not counts:
            t.counts = ctypes.POINTER(ctypes.c_uint32)()
        pal = Palette.from_buffer_copy(palette)
        rid = ctypes.c_uint64()
        _check(lib().mdzg_server_submit_image(self._s, ctypes.byref(t), aa, ctypes.byref(pal),
                                              rgb.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32)),
                                              ctypes.byref(rid)), "submit_image")
        self._live[rid.value] = (t, ((rgb, out) if counts else rgb), view, pal)
        return rid.value

    def poll(self, rid):
        return _check(lib().mdzg_server_poll(self._s, rid), "poll") == 1

    def wait(self, rid):
        info = ReqInfo()
        rc = lib().mdzg_server_wait(self._s, rid, ctypes.byref(info))
        _check(rc, "wait")
        out = self._live.pop(rid)[1]
        return out, info.as_dict()

    def info(self, rid):
        info = ReqInfo()
        _check(lib().mdzg_server_info(self._s, rid, ctypes.byref(info)), "info")
        return info.as_dict()

    def stats(self):
        st = ServeStats()
        _check(lib().mdzg_server_stats(self._s, ctypes.byref(st)), "stats")
        return st.as_dict()

    def close(self):
        if self._s:
            lib().mdzg_server_destroy(self._s)
            self._s = ctypes.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
