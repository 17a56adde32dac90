"""Generates log2_table.h: constants of the canonical FP64 log2.

log2(x) = e + log2(c_j) + P(r),  x = 2^e * m,  c_j = 1 + j/128,
r = fma(m, inv_c_j, -1),  P(r) = sum_{k=1..9} (-1)^(k+1) r^k / (k ln 2).
Every constant is computed in 60-digit decimal arithmetic and rounded once to
the nearest double (float(Decimal) rounds correctly).
"""
from decimal import Decimal, getcontext
import os

getcontext().prec = 60


def ln(x: Decimal) -> Decimal:
    return x.ln()


def main() -> None:
    ln2 = ln(Decimal(2))
    inv, lg = [], []
    for j in range(128):
        c = Decimal(1) + Decimal(j) / Decimal(128)
        inv.append(float(Decimal(1) / c))
        lg.append(float(ln(c) / ln2))
    poly = []
    for k in range(1, 10):
        a = Decimal(1) / (Decimal(k) * ln2)
        poly.append(float(a if k % 2 else -a))
    out = ["// Generated by gen_log2_table.py -- constants of the canonical FP64 log2.",
           "#pragma once",
           "#define GSAP_LOG2_INV_VALUES \\"]
    out += ["  %s, \\" % x.hex() for x in inv] + ["", "#define GSAP_LOG2_LG_VALUES \\"]
    out += ["  %s, \\" % x.hex() for x in lg] + ["", "#define GSAP_LOG2_POLY_VALUES \\"]
    out += ["  %s, \\" % x.hex() for x in poly] + ["", ""]
    here = os.path.dirname(os.path.abspath(__file__))
    with open(os.path.join(here, "log2_table.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
