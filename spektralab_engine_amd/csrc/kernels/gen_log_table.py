"""Writes spk_log_table.h: the 128-entry table of the sampler's logarithm (spk_math.h).

Entry i covers mantissas m in [1 + i/128, 1 + (i+1)/128).  invc[i] is the binary64 value nearest
1 / c_i with c_i = 1 + (i + 1/2)/128, and logc[i] is -ln(invc[i]) rounded to binary64 from a
60-digit decimal evaluation.  Entry 0 is {1, 0} (c_0 = 1), so that ln 1 = 0 exactly and inputs
just above 1 keep full relative accuracy; its |r| stays below 2^-7.  The table depends on nothing
but IEEE rounding, and the literals are exact hexadecimal floats.

    python gen_log_table.py > spk_log_table.h
"""
from decimal import Decimal, getcontext
from fractions import Fraction

getcontext().prec = 60


def main():
    rows = []
    for i in range(128):
        c = Fraction(256 + 2 * i + 1, 256) if i else Fraction(1)
        invc = float(1 / c)  # correctly rounded Fraction -> float
        exact = -(Decimal(Fraction(invc).numerator) / Decimal(Fraction(invc).denominator)).ln()
        logc = float(exact)
        rows.append((invc.hex(), logc.hex()))
    print("// Generated by gen_log_table.py: {1/c_i, -ln(1/c_i)}, c_0 = 1, c_i = 1 + (i + 1/2)/128.  Do not edit.")
    print("#pragma once")
    print("#define SPKM_LOG_TABLE_ENTRIES \\")
    for k, (a, b) in enumerate(rows):
        end = ", \\" if k < 127 else ""
        print(f"    {a}, {b}{end}")


if __name__ == "__main__":
    main()
