"""Op-level device check of the round-to-nearest arithmetic: the kernels'
own add_rn / mul_rn (mdz_amd/csrc/mp_rn.h) run on the GPU, one operand pair
per lane, against MPFR_RNDN through the oracle.  Random and edge operands mix
within every wave, so the wave-voted steps run with some lanes needing them
and some not; operands with only their top and bottom words set make the
short products fall back to the full product.  Tolerance: none."""
import ctypes
import os
import random

import numpy as np
import pytest

from oracle import oracle as O
from tests import rn_ref
from tests.gpu_arith import build as AB

pytestmark = pytest.mark.gpu

CLASSES = (1, 2, 3, 4, 6, 8, 12, 16, 24, 32)
EZERO = -(1 << 30)
VS = lambda n: ((n + 2) + 3) & ~3   # noqa: E731  (stored value size, mdzg_dev.h)


def _lib():
    if not os.path.exists(AB.LIB):
        raise RuntimeError(f"{AB.LIB} is missing: run __graft_entry__.build()")
    L = ctypes.CDLL(AB.LIB)
    u32p = ctypes.POINTER(ctypes.c_uint32)
    L.arith_check.argtypes = [ctypes.c_int] * 4 + [u32p] * 3
    L.arith_check.restype = ctypes.c_int
    return L


def _operand(rng, n, p, emin, emax):
    kind = rng.randrange(16)
    m = [rng.getrandbits(32) for _ in range(n)]
    if kind == 0:                               # power of two
        m = [0] * n
    elif kind == 1:                             # all ones
        m = [0xffffffff] * n
    elif kind == 2:                             # a few bits
        m = [(1 << rng.randrange(32)) if rng.randrange(4) == 0 else 0 for _ in range(n)]
    elif kind in (4, 5):                        # top and bottom words only
        m = [m[0]] + [0] * (n - 2) + [m[-1]] if n > 1 else m
        if kind == 5:
            m[-1] = 0x80000000
    m[-1] |= 0x80000000
    kclr = 32 * n - p
    M = sum(w << (32 * j) for j, w in enumerate(m)) >> kclr << kclr
    if kind == 3:
        return (0, EZERO, rng.getrandbits(1))
    if kind == 6 and kclr % 2 == 0:
        # 2^(32n-1) + 2^(16n + kclr/2 - 1): its square is an exact tie at
        # precision p with an even last bit (rounds down; a guessed sticky
        # bit would round up)
        M = (1 << (32 * n - 1)) + (1 << (16 * n + kclr // 2 - 1))
    return (M, rng.randint(emin, emax), rng.getrandbits(1))


def _near(rng, a, n, p):
    M, e, _ = a
    if M == 0:
        return _operand(rng, n, p, -6, 6)
    kclr = 32 * n - p
    for _ in range(3):
        bit = kclr + rng.randrange(min(p, 70))
        if bit < 32 * n - 1:
            M ^= 1 << bit
    return (M, e if rng.randrange(4) else e - 1, rng.getrandbits(1))


def _store(vals, n):
    out = np.zeros((len(vals), VS(n)), np.uint32)
    for i, (M, e, s) in enumerate(vals):
        for j in range(n):
            out[i, j] = (M >> (32 * j)) & 0xffffffff
        out[i, n] = e & 0xffffffff
        out[i, n + 1] = s
    return out


def _value(row, n):
    M = sum(int(row[j]) << (32 * j) for j in range(n))
    e = int(np.int32(row[n]))
    if M == 0 or e == EZERO:
        return rn_ref.Fraction(0)
    v = rn_ref.Fraction(M) * rn_ref.Fraction(2) ** (e - 32 * n)
    return -v if row[n + 1] else v


@pytest.mark.parametrize("n", CLASSES)
def test_ops_match_mpfr_on_device(n):
    L = _lib()
    rng = random.Random(1000 + n)
    count = 320                                  # five waves per launch
    for p in sorted({32 * n, max(2, 32 * n - 13), 32 * (n - 1) + 1 if n > 1 else 17}):
        a, b = [], []
        for i in range(count):
            x = _operand(rng, n, p, -6, 6)
            mode = rng.randrange(4)
            if mode == 0:
                y = _near(rng, x, n, p)
            else:
                spread = (4, 32 * (n + 3), 3000)[mode - 1]
                ec = 0 if x[0] == 0 else x[1]
                y = _operand(rng, n, p, ec - spread, ec + spread)
            a.append(x)
            b.append(y)
        A, B = _store(a, n), _store(b, n)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        for op in range(5):
            # op 4: |a| - |b| through the specialisation for nonnegative operands
            xa = [(M, e, 0) for M, e, _ in a] if op == 4 else a
            xb = [(M, e, 0) for M, e, _ in b] if op == 4 else b
            XA, XB = (_store(xa, n), _store(xb, n)) if op == 4 else (A, B)
            R = np.zeros_like(A)
            rc = L.arith_check(n, op, 32 * n - p, count, XA.ctypes.data_as(u32p),
                               XB.ctypes.data_as(u32p), R.ctypes.data_as(u32p))
            assert rc == 0, rc
            for i in range(count):
                sa = _to_str(xa[i], n)
                sb = _to_str(xa[i] if op == 3 else xb[i], n)
                want = rn_ref.parse(O.rn_op((0, 1, 2, 2, 1)[op], p, sa, sb))
                got = _value(R[i], n)
                assert got == want, (n, p, op, i, sa, sb)


def _to_str(v, n):
    """Exact hex-float string of a stored value (significand M, exponent e)."""
    M, e, s = v
    if M == 0 or e == EZERO:
        return "0"
    return ("-" if s else "") + hex(M) + "p" + str(e - 32 * n)
