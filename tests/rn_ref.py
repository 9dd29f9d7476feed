"""Round-to-nearest-even arithmetic on exact rationals (pure Python).

An independent restatement of what MPFR_RNDN fixes by definition: the exact
result of each operation, rounded once to p significant bits, ties to even.
Used to pin the oracle's MPFR loops, its c chain (src/fractal.c:159-188) and
its rect chain (src/coords.c:265-313) on small cases.
"""
from fractions import Fraction


def parse(s):
    """Exact value of a decimal or C99/MPFR hex-float string."""
    s = s.strip()
    neg = s.startswith("-")
    if s[0] in "+-":
        s = s[1:]
    if s.lower().startswith("0x"):
        body = s[2:]
        if "p" in body.lower():
            i = body.lower().index("p")
            mant, ex = body[:i], int(body[i + 1:])
        else:
            mant, ex = body, 0
        if "." in mant:
            ip, fp = mant.split(".")
        else:
            ip, fp = mant, ""
        m = int((ip or "0") + fp, 16)
        v = Fraction(m) * Fraction(2) ** (ex - 4 * len(fp))
    else:
        v = Fraction(s)
    return -v if neg else v


def ilog2(v):
    """floor(log2(v)) for a positive Fraction."""
    n, d = v.numerator, v.denominator
    e = n.bit_length() - d.bit_length()
    # adjust so that 2^e <= v < 2^(e+1)
    if e >= 0:
        if n < (d << e):
            e -= 1
    else:
        if (n << -e) < d:
            e -= 1
    return e


def rn(v, p):
    """Round v (Fraction) to p significant bits, ties to even."""
    v = Fraction(v)
    if v == 0:
        return Fraction(0)
    neg = v < 0
    a = -v if neg else v
    e = ilog2(a)
    scale = Fraction(2) ** (e + 1 - p)
    m = a / scale                      # in [2^(p-1), 2^p)
    q = m.numerator // m.denominator
    r = m - q
    if r > Fraction(1, 2) or (r == Fraction(1, 2) and q % 2 == 1):
        q += 1
    out = q * scale
    return -out if neg else out


def formula(fractal, depth, p, wim, wre, c_im, c_re, wim2, wre2):
    """The MPFR loop of src/frac_mandel.c:33-51 and siblings, in exact RN arithmetic."""
    for wz in range(1, depth + 1):
        t = rn(wre * wim, p)
        if fractal == 1:
            t = abs(t)
        t = 2 * t
        wim = rn(t + c_im, p)
        t = rn(wre2 - wim2, p)
        if fractal == 2 or (fractal == 3 and wz % 2):
            t = abs(t)
        wre = rn(t + c_re, p)
        wim2 = rn(wim * wim, p)
        wre2 = rn(wre * wre, p)
        if rn(wim2 + wre2, p) > 4:
            return wz
    return 0


def gmp_mpf_bits(prec):
    """mpf_get_prec after mpf_init2(prec) for 64-bit limbs (GMP's documented rule:
    limbs = (max(53, prec) + 2*64 - 1) // 64, bits = (limbs - 1) * 64)."""
    limbs = (max(53, prec) + 127) // 64
    return (limbs - 1) * 64


def pixel_c(xmin, ymax, width, p, W, line, ix):
    """src/fractal.c:159-188 for one pixel (Fractions at the view's precision)."""
    img_rw = rn(W, p)
    img_xmin = rn(xmin, p)
    wd = rn(width, p)
    t1 = rn(wd / img_rw, p)
    t1 = rn(t1 * line, p)
    y = rn(ymax - t1, p)
    t1 = rn(Fraction(ix) / img_rw, p)
    x = rn(t1 * wd, p)
    x = rn(x + img_xmin, p)
    return x, y


def rect_from_center(cx, cy, size, p, user_w, user_h, coords=False):
    """src/coords.c:204-231 + :265-313 + coords_get_rect, exact RN arithmetic;
    coords=True: the values at the coords precision (coords_get_rect_gmp)."""
    cp = max(p, 80)
    cp = max(cp, gmp_mpf_bits(cp))
    hp = max(p, 80)
    aspect = Fraction(user_w / user_h)          # the C double (double)w/h
    cx, cy, size = rn(parse(cx), p), rn(parse(cy), p), rn(parse(size), p)
    if user_w / user_h > 1.0:
        width = size
        height = rn(width / aspect, cp)
        tmp = rn(width / 2, cp)
        xmin = rn(cx - tmp, cp)
        xmax = rn(xmin + width, cp)
        tmp = rn(tmp / aspect, cp)
        ymin = rn(cy - tmp, cp)
        ymax = rn(ymin + height, cp)
    else:
        height = size
        width = rn(height * aspect, cp)
        tmp = rn(height / 2, cp)
        ymin = rn(cy - tmp, cp)
        ymax = rn(ymin + height, cp)
        tmp = rn(tmp * aspect, cp)
        xmin = rn(cx - tmp, cp)
        xmax = rn(xmin + width, cp)
    if coords:
        return (xmin, xmax, ymax, width)
    return tuple(rn(v, hp) for v in (xmin, xmax, ymax, width))


def rect_from_corners(xmin, xmax, ymax, p, user_w, user_h, coords=False, file_aspect=None):
    """A corner-form settings file's rect at user_w x user_h, exact RN
    arithmetic at the coords precision (coords=True keeps that precision).

    file_aspect given: as MDZ's command line renders it -- the centre from
    coords_set_rect / coords_rect_to_center (src/coords.c:327-333, :234-252)
    at the aspect of the 480-wide load, 480 / int(480 / file_aspect), then the
    stale size 4.0 that init_misc's coords_set restores (:255-262) as the width
    (wide) or height (tall), then coords_center_to_rect (:265-313).
    file_aspect None: the file's rect refitted at the output aspect."""
    cp = max(p, 80)
    cp = max(cp, gmp_mpf_bits(cp))
    hp = max(p, 80)
    ad = user_w / user_h
    aspect = Fraction(ad)
    aspect0 = aspect if file_aspect is None else Fraction(480 / int(480 / file_aspect))
    xmin, xmax, ymax = (rn(rn(parse(s), p), cp) for s in (xmin, xmax, ymax))
    width = rn(xmax - xmin, cp)
    height = rn(width / aspect0, cp)
    ymin = rn(ymax - height, cp)
    cx = rn(rn(xmin + xmax, cp) / 2, cp)
    cy = rn(rn(ymin + ymax, cp) / 2, cp)
    if file_aspect is not None:
        if ad > 1.0:
            width = Fraction(4)
        else:
            height = Fraction(4)
    if ad > 1.0:
        height = rn(width / aspect, cp)
        tmp = rn(width / 2, cp)
        xmin = rn(cx - tmp, cp)
        xmax = rn(xmin + width, cp)
        tmp = rn(tmp / aspect, cp)
        ymin = rn(cy - tmp, cp)
        ymax = rn(ymin + height, cp)
    else:
        width = rn(height * aspect, cp)
        tmp = rn(height / 2, cp)
        ymin = rn(cy - tmp, cp)
        ymax = rn(ymin + height, cp)
        tmp = rn(tmp * aspect, cp)
        xmin = rn(cx - tmp, cp)
        xmax = rn(xmin + width, cp)
    if coords:
        return (xmin, xmax, ymax, width)
    return tuple(rn(v, hp) for v in (xmin, xmax, ymax, width))
