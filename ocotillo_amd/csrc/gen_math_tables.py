"""Generates oco_math_tables.h, the constant tables of oco_math.h.

Every entry is computed here from its mathematical definition with Python's
decimal module at 60 significant digits and rounded once to binary64, so the
header can be regenerated and checked at any time (tests compare it with a
fresh run of this script).

Tables
  exp:  2^(i/128) = THI[i] * (1 + TLO[i]),  THI = RN(2^(i/128)),
        TLO = RN(2^(i/128)/THI - 1),                          i = 0..127
  log:  the mantissa range [0x1.6ap-1, 0x1.6ap0) is cut into 128 cells by
        the top 7 mantissa bits of (bits(x) - bits(0x1.6ap-1)).  For cell i,
        INVC[i] is a multiple of 2^-7 (cells below 1) or 2^-8 (cells above 1)
        chosen so that |z*INVC - 1| < 2^-7 on the whole cell; then
        r = fma(z, INVC, -1) is exact.  LOGC = -log(INVC) as a double-double.
"""
import os
import sys
from decimal import Decimal, getcontext
from fractions import Fraction
import struct

getcontext().prec = 60
OFF = 0x3FE6A00000000000


def as_double(u):
    return struct.unpack("<d", struct.pack("<Q", u))[0]


def as_bits(d):
    return struct.unpack("<Q", struct.pack("<d", d))[0]


def rn(dec):
    """Round a Decimal to the nearest binary64 (ties to even)."""
    return float(dec)


def dd(dec):
    hi = rn(dec)
    lo = rn(dec - Decimal(hi))
    return hi, lo


def exp_tables():
    thi, tlo = [], []
    ln2 = Decimal(2).ln()
    for i in range(128):
        v = (ln2 * i / 128).exp()
        h = rn(v)
        thi.append(h)
        tlo.append(rn(v / Decimal(h) - 1))
    return thi, tlo


def log_tables():
    invc, lhi, llo = [], [], []
    for i in range(128):
        z0 = Fraction(as_double(OFF + (i << 45)))
        z1 = Fraction(as_double(OFF + ((i + 1) << 45)))
        # cells below 1 take invc = j/128 (>= 1), cells at/above 1 take j/256
        s = 128 if z1 <= 1 else 256
        best = None
        for j in range(int(s * 0.5), int(s * 2) + 1):
            ic = Fraction(j, s)
            err = max(abs(z0 * ic - 1), abs(z1 * ic - 1))
            if best is None or err < best[0]:
                best = (err, ic)
        err, ic = best
        # the two cells touching 1.0 use invc = 1 exactly: then log c = 0 and
        # log x = log1p(x - 1) with x - 1 exact, accurate to the last bit near 1
        if z1 == 1 or z0 == 1:
            ic = Fraction(1)
            err = max(abs(z0 - 1), abs(z1 - 1))
        # cells are half-open [z0, z1): the bound at z1 is not attained
        assert err <= Fraction(1, 128), (i, float(err))
        assert (z1 <= 1) == (ic >= 1) or ic == 1, i
        c = float(ic)
        assert Fraction(c) == ic
        v = -Decimal(ic.numerator).ln() + Decimal(ic.denominator).ln()
        h, l = dd(v)
        invc.append(c)
        lhi.append(h)
        llo.append(l)
    return invc, lhi, llo


def fmt(name, vals):
    body = ", \\\n".join("    " + ", ".join(v.hex() for v in vals[k:k + 4]) for k in range(0, len(vals), 4))
    return f"#define {name} {{ \\\n{body} \\\n}}\n"


def render():
    thi, tlo = exp_tables()
    invc, lhi, llo = log_tables()
    out = ["/* Generated by gen_math_tables.py -- do not edit. */\n",
           "#ifndef OCO_MATH_TABLES_H\n#define OCO_MATH_TABLES_H\n",
           fmt("OCO_EXP_THI_INIT", thi), fmt("OCO_EXP_TLO_INIT", tlo),
           fmt("OCO_LOG_INVC_INIT", invc), fmt("OCO_LOG_LHI_INIT", lhi), fmt("OCO_LOG_LLO_INIT", llo),
           "#endif\n"]
    return "".join(out)


if __name__ == "__main__":
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "oco_math_tables.h")
    text = render()
    if len(sys.argv) > 1 and sys.argv[1] == "--check":
        sys.exit(0 if open(path).read() == text else 1)
    with open(path, "w") as f:
        f.write(text)
