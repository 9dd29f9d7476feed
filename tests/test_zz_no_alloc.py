"""mdzg_render_tile_noalloc, the entry point that allocates no device memory,
against mdzg_render_tile on the same tiles: identical counts (uint32, bit for
bit), two calls in a row with the second on a smaller tile, on every
arithmetic (FP64, MPFR at 128 and 1024 bits, long double, GMP).  The counts
buffers are device memory the test allocates itself, through the HIP runtime
libmdzg.so runs on."""
import ctypes

import numpy as np
import pytest

import mdz_amd as M

pytestmark = pytest.mark.gpu

# a gallery view (src/gallery subdudehue's neighbourhood) with escape structure
CX, CY, SIZE = "-1.7495931476527795538581440", "5.79069164258e-8", "3e-9"


def _view(precision, w, h, path=M.PATH_MPFR, depth=500, fractal=M.MANDELBROT):
    rect = M.rect_from_center(CX, CY, SIZE, precision, w, h, coords=(path == M.PATH_GMP))
    return M.make_view(rect, precision=precision, depth=depth, width=w, height=h, path=path,
                       fractal=fractal)


_HIP = None


def hip():
    """The HIP runtime libmdzg.so loaded (by its soname, already in the process)."""
    global _HIP
    if _HIP is None:
        M.lib()
        L = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        L.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        L.hipFree.argtypes = [vp]
        L.hipMemset.argtypes = [vp, ctypes.c_int, ctypes.c_size_t]
        L.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        L.hipDeviceSynchronize.argtypes = []
        _HIP = L
    return _HIP


class DevBuf:
    """nwords uint32 of device memory, filled with the byte `fill`."""

    def __init__(self, nwords, fill=0x5a):
        self.n = nwords
        self.p = ctypes.c_void_p()
        assert hip().hipMalloc(ctypes.byref(self.p), 4 * nwords) == 0
        assert hip().hipMemset(self.p, fill, 4 * nwords) == 0
        assert hip().hipDeviceSynchronize() == 0

    def read(self):
        a = np.empty(self.n, dtype=np.uint32)
        assert hip().hipMemcpy(a.ctypes.data, self.p, 4 * self.n, 2) == 0   # device to host
        return a

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


def _noalloc(ctx, view, lines=None, cols=None):
    nl = lines[1] if lines else view.real_height
    nc = cols[1] if cols else view.real_width
    buf = DevBuf(nl * nc)
    st = ctx.render_noalloc(view, buf.p.value, lines, cols)
    return buf.read().reshape(nl, nc), st


CASES = [
    ("mpfr128 then mpfr1024", (_view, (128, 96, 72), {}, None, None),
     (_view, (1024, 48, 36), {"depth": 300}, (5, 20), (3, 40))),
    ("fp64 then ld", (_view, (64, 120, 90), {"path": M.PATH_DOUBLE, "depth": 2000}, None, None),
     (_view, (64, 64, 48), {"path": M.PATH_LD, "depth": 2000}, (10, 8), (0, 64))),
    ("gmp256 then mpfr176 celtic",
     (_view, (256, 64, 48), {"path": M.PATH_GMP}, None, None),
     (_view, (176, 40, 30), {"fractal": M.GENERALIZED_CELTIC}, (2, 20), (5, 30))),
]


@pytest.mark.parametrize("name,first,second", CASES, ids=[c[0] for c in CASES])
def test_noalloc_equals_render_tile_twice(ctx, name, first, second):
    distinct = []
    for mk, args, kw, lines, cols in (first, second):
        v = mk(*args, **kw)
        print(f"noalloc case {name}: p={v.precision} path={v.path} "
              f"{v.real_width}x{v.real_height} lines={lines} cols={cols}", flush=True)
        want, _ = ctx.render(v, lines=lines, cols=cols)
        got, st = _noalloc(ctx, v, lines, cols)
        assert np.array_equal(got, want), name
        assert st["launches"] == 1 and st["kernel_ms"] > 0.0
        distinct.append(len(np.unique(want)))
    # the maps carry escape structure: a kernel that never escaped would fail
    assert max(distinct) > 5, distinct


def test_noalloc_rejects_what_it_cannot_take(ctx):
    # start values beyond the static array: (ncols + nlines + 2) values of
    # 144 bytes at 1024 bits exceed 8 MiB from 58251 values on
    v = _view(1024, 60000, 1, depth=10)
    buf = DevBuf(60000, fill=0)
    with pytest.raises(M.MdzgError) as e:
        ctx.render_noalloc(v, buf.p.value)
    assert e.value.code == -7
    assert not buf.read().any()                    # nothing was written
    # host memory is not a device counts buffer
    host = np.zeros(96 * 72, dtype=np.uint32)
    with pytest.raises(M.MdzgError) as e:
        ctx.render_noalloc(_view(128, 96, 72), host.ctypes.data)
    assert e.value.code == -1
