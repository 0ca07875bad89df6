"""Generate cyl_crmath_tables.h: double-double constants of the correctly
rounded exp/log used on the device (exact rationals and 60-digit decimal
arithmetic, each value split as hi = nearest double, lo = nearest double of
the remainder).

    python cylag_amd/csrc/gen_crmath_tables.py > cylag_amd/csrc/cyl_crmath_tables.h
"""
from decimal import Decimal, getcontext
from fractions import Fraction
from math import factorial

getcontext().prec = 60


def split(v):
    """v (Fraction or Decimal) -> (hi, lo) doubles with hi + lo ~ v to ~2^-107."""
    if isinstance(v, Fraction):
        hi = float(v)
        lo = float(v - Fraction(hi))
    else:
        hi = float(v)
        lo = float(v - Decimal(hi))
    return hi, lo


def fmt(pair):
    return "{%s, %s}" % (pair[0].hex(), pair[1].hex())


def main():
    out = []
    out.append("// Generated by gen_crmath_tables.py -- do not edit.")
    out.append("#pragma once")
    out.append("namespace cyl {")
    ln2 = Decimal(2).ln()
    hi, lo = split(ln2)
    out.append("constexpr double CR_LN2_HI = %s, CR_LN2_LO = %s;" % (hi.hex(), lo.hex()))
    out.append("constexpr double CR_INV_LN2 = %s;" % float(1 / ln2).hex())
    # 1/n! for the exp polynomial on |r| <= 1/128 (n = 0..12) and expm1 on |x| < 1/32 (n = 0..15)
    out.append("constexpr int CR_NFACT = 16;")
    out.append("__constant__ DD CR_INVFACT[CR_NFACT] = {")
    out.append(",\n".join("  " + fmt(split(Fraction(1, factorial(n)))) for n in range(16)))
    out.append("};")
    # e^(j/64), j = -32..32
    out.append("__constant__ DD CR_EXPT[65] = {")
    out.append(",\n".join("  " + fmt(split((Decimal(j) / 64).exp())) for j in range(-32, 33)))
    out.append("};")
    # e^(j/256), j = -96..96 (the fast paths: |r| <= ln2/2 gives |j| <= 89)
    out.append("constexpr int CR_NT256 = 96;")
    out.append("__constant__ DD CR_EXPT256[2 * CR_NT256 + 1] = {")
    out.append(",\n".join("  " + fmt(split((Decimal(j) / 256).exp())) for j in range(-96, 97)))
    out.append("};")
    # 2/(2n+1) for log1p(t) = 2 atanh(u), u = t/(2+t), |u| < 1/64 (n = 0..11)
    out.append("constexpr int CR_NATANH = 12;")
    out.append("__constant__ DD CR_ATANH[CR_NATANH] = {")
    out.append(",\n".join("  " + fmt(split(Fraction(2, 2 * n + 1))) for n in range(12)))
    out.append("};")
    out.append("}  // namespace cyl")
    print("\n".join(out))


if __name__ == "__main__":
    main()
