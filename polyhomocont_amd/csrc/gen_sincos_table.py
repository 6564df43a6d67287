"""Writes phc_sincos_table.h: sin(i/128) and cos(i/128), i = 0..127, each as
an unevaluated sum hi + lo of two doubles (hi the nearest double, lo the
nearest double to the remainder), the table of glibc's dbl-64 sin/cos
(sysdeps/ieee754/dbl-64/sincostab.c holds the same values).

Run once by hand when the header has to be regenerated:
    python polyhomocont_amd/csrc/gen_sincos_table.py
The header is committed; the build does not run this script. The values
are pinned by tests/test_math_host.py, which checks libm_sincos against the
host's sincos bit for bit.
"""
import os

import mpmath

HERE = os.path.dirname(os.path.abspath(__file__))


def split(v):
    hi = float(v)
    return hi, float(v - mpmath.mpf(hi))


def main():
    mpmath.mp.prec = 300
    rows = []
    for i in range(128):
        x = mpmath.mpf(i) / 128
        sh, sl = split(mpmath.sin(x))
        ch, cl = split(mpmath.cos(x))
        rows.append(f"    {sh.hex()}, {sl.hex()}, {ch.hex()}, {cl.hex()},  // {i}/128")
    text = (
        "// Generated by gen_sincos_table.py: {sin hi, sin lo, cos hi, cos lo}\n"
        "// of i/128, i = 0..127, as double-double values -- the table of the IBM\n"
        "// Accurate Mathematical Library sin/cos method of glibc 2.35\n"
        "// (sysdeps/ieee754/dbl-64, LGPL-2.1-or-later), computed here from its\n"
        "// definition (mpmath, 300 bits), not copied.\n"
        "#pragma once\n\n"
        "namespace phc {\n\n"
        "constexpr double SINCOS_TABLE[512] = {\n" + "\n".join(rows) + "\n};\n\n"
        "}  // namespace phc\n")
    with open(os.path.join(HERE, "phc_sincos_table.h"), "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
