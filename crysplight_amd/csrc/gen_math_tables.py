"""Writes math_tables.h, the two tables of the transcendentals in crysp_device.h:

SINCOS_TAB[k] = {sin(k pi/128), cos(k pi/128)}, k = 0..256 (f_sincos);
LOG_TAB[i]    = {1/c, log(c)} for the 129 intervals of m in [sqrt(2)/2, sqrt(2)) picked by
                the low exponent bit and the 7 leading mantissa bits of m (f_log): c is the
                interval's midpoint, except 1 for the two intervals next to 1; the first
                column is 1/c rounded to binary64, the second log of the exact inverse of
                that rounded value, so that m (1/c) - 1 and the logarithm's decomposition
                agree.

Every value is computed in decimal arithmetic with 60 significant digits (pi by Machin's
formula, sin and cos by their Taylor series, log through the atanh series) and rounded once
to binary64 by Python's correctly rounded decimal-to-float conversion; multiples of pi/2 get
their exact sin and cos.  Run: python crysplight_amd/csrc/gen_math_tables.py
"""
from decimal import Decimal, getcontext
from pathlib import Path

getcontext().prec = 60
EPS = Decimal(10) ** -58


def _atan_inv(n: int) -> Decimal:
    x = Decimal(1) / n
    x2, s, t, k = x * x, x, x, 1
    while True:
        t *= -x2
        k += 2
        term = t / k
        if abs(term) < EPS:
            return s
        s += term


PI = 16 * _atan_inv(5) - 4 * _atan_inv(239)


def sincos(a: Decimal):
    s, c, t, n = Decimal(0), Decimal(0), Decimal(1), 0
    while abs(t) >= EPS:
        if n % 4 == 0:
            c += t
        elif n % 4 == 1:
            s += t
        elif n % 4 == 2:
            c -= t
        else:
            s -= t
        n += 1
        t = t * a / n
    return s, c


def ln(x: Decimal) -> Decimal:
    """log x = 2 atanh((x - 1) / (x + 1))."""
    s = (x - 1) / (x + 1)
    s2, t, acc, k = s * s, s, Decimal(0), 1
    while True:
        term = t / k
        if abs(term) < EPS:
            return 2 * acc
        acc += term
        t *= s2
        k += 2


def sincos_table():
    rows = []
    for k in range(257):
        if k % 64 == 0:
            q = k // 64
            rows.append(((0.0, 1.0, 0.0, -1.0, 0.0)[q], (1.0, 0.0, -1.0, 0.0, 1.0)[q]))
        else:
            s, c = sincos(PI * k / 128)
            rows.append((float(s), float(c)))
    return rows


LOG_FIRST = 0x35          # (0x3F3504F3 >> 16) & 0xFF: m = sqrt(2)/2
LOG_ONE = (0x7F, 0x80)    # the intervals [255/256, 1) and [1, 129/128)


def log_table():
    rows = []
    for idx in range(LOG_FIRST, 0xB6):
        if idx in LOG_ONE:
            rows.append((1.0, 0.0))
            continue
        j = idx & 127
        scale = 256 if idx < 128 else 128            # m < 1: exponent -1; m >= 1: exponent 0
        c = (Decimal(128 + j) + Decimal("0.5")) / scale
        inv = float(1 / c)
        rows.append((inv, float(-ln(Decimal(inv)))))
    return rows


def _rows(name, n, rows):
    out = [f"__constant__ __attribute__((aligned(16))) double {name}[{n}][2] = {{"]
    out += [f"    {{{a.hex()}, {b.hex()}}},  // {k}" for k, (a, b) in enumerate(rows)]
    out.append("};")
    return out


def header() -> str:
    out = ["// math_tables.h -- written by gen_math_tables.py; do not edit.",
           "// SINCOS_TAB[k] = {sin(k pi/128), cos(k pi/128)}, k = 0..256, correctly rounded binary64.",
           "// LOG_TAB[i] = {1/c, log(c)} for the interval of m whose low exponent bit and 7 leading",
           "// mantissa bits are 0x35 + i (i = 0..128), 1/c rounded and log(c) = -log(1/c) of it.",
           "#pragma once", ""]
    out += _rows("SINCOS_TAB", 257, sincos_table())
    out.append("")
    out += _rows("LOG_TAB", 129, log_table())
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    p = Path(__file__).resolve().parent / "math_tables.h"
    p.write_text(header())
    print(p)
