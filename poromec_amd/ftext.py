"""Text as the reference's Fortran output statements write it: the edit
descriptors Aw, Ew.d, Ew.dEe, Fw.d and Iw, and list-directed output of integers
and reals as amdflang's runtime writes them.

Ew.d[Ee] writes 0.d1..dd and an exponent; without Ee the exponent is E+xx
and drops its letter at three digits (+xxx), with Ee it is E+ and e digits.
d significant digits, rounded to nearest from the binary value.

List-directed output starts a record with one blank.  Character items follow
each other with no separator; a number after a character item is preceded
by one blank.  A real uses the fewest digits that read back to the same
value.  It is written in F form when its one-digit rounding 0.d x 10**x has
0 <= x <= 15, with no leading zero (``.25``, ``3600.``).  Otherwise it is
written in 1P E form with at least two exponent digits (``1.E-05``,
``2.777777777777778E-04``, ``1.E+100``).

The Fortran standard leaves list-directed digits to the processor.  Up to 16
significant digits the choice below is amdflang's: the exact value cut to
that length when the cut reads back, else the cut plus one unit.  Where 17
digits are needed amdflang sometimes takes another of the several 17-digit
strings that read back; both read back to the same value.  Subnormal
values are not covered."""
from __future__ import annotations

import math
from decimal import ROUND_DOWN, Decimal, localcontext

import numpy as np


def _special(v: float, w: int):
    if v != v:
        return "NaN".rjust(w)
    if math.isinf(v):
        return ("Infinity" if v > 0 else "-Infinity").rjust(w)
    return None


def fortran_e(v: float, w: int = 18, d: int = 10, e: int | None = None) -> str:
    """``v`` under the edit descriptor Ew.d (e None) or Ew.dEe."""
    v = float(v)
    sp = _special(v, w)
    if sp is not None:
        return sp
    if v == 0.0:
        sign, digits, exp = ("-" if np.signbit(v) else ""), "0" * d, 0
    else:
        sign = "-" if v < 0 else ""
        m, x = f"{abs(v):.{d - 1}e}".split("e")
        digits, exp = m.replace(".", ""), int(x) + 1
    es = "+" if exp >= 0 else "-"
    if e is None:
        tail = f"E{es}{abs(exp):02d}" if abs(exp) <= 99 else f"{es}{abs(exp):03d}"
    else:
        if abs(exp) >= 10 ** e:
            return "*" * w
        tail = f"E{es}{abs(exp):0{e}d}"
    s = sign + "0." + digits + tail
    if len(s) > w:  # the optional leading zero goes first
        s = sign + "." + digits + tail
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_f(v: float, w: int, d: int) -> str:
    """``v`` under Fw.d."""
    v = float(v)
    sp = _special(v, w)
    if sp is not None:
        return sp
    s = f"{v:.{d}f}"
    if len(s) > w and s.lstrip("-").startswith("0."):
        s = s.replace("0.", ".", 1)
    return s.rjust(w) if len(s) <= w else "*" * w


def fortran_a(s: str, w: int) -> str:
    """``s`` under Aw: its first w characters, or right-justified when
    shorter."""
    return s[:w] if len(s) >= w else s.rjust(w)


def fortran_i(v: int, w: int) -> str:
    """``v`` under Iw."""
    s = str(int(v))
    return s.rjust(w) if len(s) <= w else "*" * w


def list_real(v: float) -> str:
    """A real(8) item of list-directed output, without the separating blank."""
    v = float(v)
    if v != v:
        return "NaN"
    if math.isinf(v):
        return "Inf" if v > 0 else "-Inf"
    sign = "-" if np.signbit(v) else ""
    if v == 0.0:
        return sign + "0."
    a = abs(v)
    digits, xp = _shortest(a)
    one = int(f"{a:.0e}".split("e")[1]) + 1  # exponent of the one-digit rounding
    if 0 <= one <= 15:
        if xp > 0:
            ipart = digits[:xp].ljust(xp, "0")
            fpart = digits[xp:]
        else:
            ipart = ""
            fpart = "0" * (-xp) + digits
        return sign + ipart + "." + fpart
    e1 = xp - 1
    es = "+" if e1 >= 0 else "-"
    return sign + digits[0] + "." + digits[1:] + f"E{es}{abs(e1):02d}"


def _shortest(a: float):
    """The fewest decimal digits D (a = 0.D x 10**xp) that read back to a:
    for each length, the exact value cut to that many digits if it lies in
    a's rounding interval, else that cut plus one in its last digit (so not
    always the nearer of the two)."""
    with localcontext() as ctx:
        ctx.prec = 1100  # exact for every double's neighbours' midpoints
        ex = Decimal(a)
        below = Decimal(math.nextafter(a, 0.0))
        nxt = math.nextafter(a, math.inf)
        above = Decimal(nxt) if not math.isinf(nxt) else 2 * ex - below
        lo, hi = (ex + below) / 2, (ex + above) / 2

        def inside(x):  # the interval's ends excluded, ties to even or not
            return lo < x < hi

        lead = ex.adjusted()  # a = d.ddd x 10**lead
        for n in range(1, 18):
            unit = Decimal(1).scaleb(lead - n + 1)
            cut = ex.quantize(unit, rounding=ROUND_DOWN)
            for cand in (cut, cut + unit):
                if inside(cand):
                    t = cand.as_tuple()
                    d = "".join(map(str, t.digits)).rstrip("0") or "0"
                    return d, cand.adjusted() + 1
    d = f"{a:.16e}"
    m, x = d.split("e")
    return m.replace(".", "").rstrip("0"), int(x) + 1


def list_int(v: int) -> str:
    return str(int(v))


def list_record(*items) -> str:
    """One list-directed output record: str items are character items,
    int and float items numbers."""
    out = [" "]
    for it in items:
        if isinstance(it, str):
            out.append(it)
        else:
            if len(out) > 1:
                out.append(" ")
            out.append(list_int(it) if isinstance(it, (int, np.integer)) else list_real(it))
    return "".join(out)
