"""Writes sincos_table.h: sin and cos of k*pi/128, k = 0..256, as correctly rounded binary64
values (the table f_sincos in crysp_device.h reads).

The values are computed in decimal arithmetic with 60 significant digits (pi by Machin's
formula, sin and cos by their Taylor series) and rounded once to binary64 by Python's
correctly rounded decimal-to-float conversion; the multiples of pi/2 are set to their exact
values.  Run: python crysplight_amd/csrc/gen_sincos_table.py
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


def table():
    rows = []
    for k in range(257):
        if k % 64 == 0:
            q = k // 64
            rows.append(((0.0, 1.0, 0.0, -1.0, 0.0)[q], (1.0, 0.0, -1.0, 0.0, 1.0)[q]))
        else:
            s, c = sincos(PI * k / 128)
            rows.append((float(s), float(c)))
    return rows


def header() -> str:
    out = ["// sincos_table.h -- written by gen_sincos_table.py; do not edit.",
           "// SINCOS_TAB[k] = {sin(k pi/128), cos(k pi/128)}, k = 0..256, correctly rounded binary64.",
           "#pragma once", "",
           "__constant__ __attribute__((aligned(16))) double SINCOS_TAB[257][2] = {"]
    for k, (s, c) in enumerate(table()):
        out.append(f"    {{{s.hex()}, {c.hex()}}},  // {k}")
    out.append("};")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    p = Path(__file__).resolve().parent / "sincos_table.h"
    p.write_text(header())
    print(p)
