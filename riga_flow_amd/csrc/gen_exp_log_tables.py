"""Writes exp_log_tables.h: the two tables of exp_log.h, from 60-digit decimal
arithmetic (Python's decimal exp and ln are correctly rounded).

exp: 2^(j/128) for j = 0..127 as a pair (hi = the nearest double, lo = the
nearest double to the remainder).

log: y in [0.6875, 1.375) cut into 128 cells -- 80 of width 1/256 below 1, 48 of
width 1/128 above -- the cell selected by the top 7 bits of (bits(x) - bits(0.6875)).
Per cell: invc, an 8-significant-bit approximation of 1 / (cell middle) (exactly 1
for the two cells that touch 1), and -ln(invc) as hi + lo.  With 8 bits and
|y invc - 1| < 2^-7 the product y * invc - 1 is exact in one fused multiply-add
(checked here on every cell's end points and on random points), so the reduced
argument carries no rounding error.

    python3 gen_exp_log_tables.py        # rewrites exp_log_tables.h beside this file
"""
import math
import os
import random
import struct
from decimal import Decimal, getcontext
from fractions import Fraction

getcontext().prec = 60
HERE = os.path.dirname(os.path.abspath(__file__))
EXP_N = 128
LOG_N = 128
LOG_OFF = 0x3FE6000000000000  # bits of 0.6875


def exp_rows():
    ln2 = Decimal(2).ln()
    rows = []
    for j in range(EXP_N):
        v = (ln2 * j / EXP_N).exp()
        hi = float(v)
        rows.append((hi, float(v - Decimal(hi))))
    return rows


def log_cell(j):
    if j < 80:
        a = Fraction(11, 16) + Fraction(j, 256)
        return a, a + Fraction(1, 256)
    a = 1 + Fraction(j - 80, 128)
    return a, a + Fraction(1, 128)


def log_rows():
    rows = []
    for j in range(LOG_N):
        a, b = log_cell(j)
        if j in (79, 80):  # the cells at 1: no cancellation between ln c and ln(1 + r)
            invc = 1.0
        else:
            t = float(2 / (a + b))
            q = 2.0 ** (math.frexp(t)[1] - 8)  # 8 significant bits
            cands = (math.floor(t / q) * q, math.ceil(t / q) * q)
            invc = min(cands, key=lambda c: max(abs(Fraction(c) * a - 1), abs(Fraction(c) * b - 1)))
        assert max(abs(Fraction(invc) * a - 1), abs(Fraction(invc) * b - 1)) <= Fraction(1, 128)
        lc = -Decimal(invc).ln()
        hi = float(lc)
        rows.append((invc, hi, float(lc - Decimal(hi))))
    return rows


def check_log_exact(rows):
    """y * invc - 1 is a double for every y of every cell (end points and 2,000
    random doubles per cell): the fused multiply-add computes it exactly."""
    rng = random.Random(5)
    for j, (invc, _, _) in enumerate(rows):
        a, b = log_cell(j)
        ys = [float(a), math.nextafter(float(b), 0.0)]
        ys += [float(a) + rng.random() * float(b - a) for _ in range(2000)]
        for y in ys:
            exact = Fraction(y) * Fraction(invc) - 1
            assert Fraction(float(exact)) == exact, (j, y)
            bits = struct.unpack("<Q", struct.pack("<d", y))[0]
            assert ((bits - LOG_OFF) >> 45) & 127 == j, (j, y)  # the cell the top bits select


def main():
    er, lr = exp_rows(), log_rows()
    check_log_exact(lr)
    out = ["/* exp_log_tables.h -- written by gen_exp_log_tables.py; do not edit. */",
           "",
           "/* 2^(j/128), j = 0..127: {hi, lo} */",
           "RIGA_EXP_LOG_TABLE double riga_exp_tab[%d] = {" % (2 * EXP_N)]
    out += ["    %s, %s," % (hi.hex(), lo.hex()) for hi, lo in er]
    out += ["};", "",
            "/* log cells j = 0..127: {invc, -ln(invc) hi, -ln(invc) lo, 0} */",
            "RIGA_EXP_LOG_TABLE double riga_log_tab[%d] = {" % (4 * LOG_N)]
    out += ["    %s, %s, %s, 0.0," % (c.hex(), hi.hex(), lo.hex()) for c, hi, lo in lr]
    out += ["};", ""]
    with open(os.path.join(HERE, "exp_log_tables.h"), "w") as f:
        f.write("\n".join(out))


if __name__ == "__main__":
    main()
