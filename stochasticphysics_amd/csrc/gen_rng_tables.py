"""Generate sp_rng_tables.h: the constant tables of the FP64 Box-Muller transform.

    python stochasticphysics_amd/csrc/gen_rng_tables.py     # rewrites sp_rng_tables.h beside this file

The tables are part of the noise-stream definition (sp_rng.h), so the header
is committed and the build never regenerates it.  Every entry is the double
nearest to the stated value as Python's math module (the C library)
computes it from the given double arguments:
  * SP_SINCOS_TAB[j] = (sin t_j, cos t_j), t_j = (j + 1/2) * (2 pi / 128), j = 0..127
  * SP_LOG_TAB[j + 32] = (log c_j, 2^-7 / c_j), c_j = 1 + j / 128, j = -32..64
    (2^-7 / c_j is the double nearest to 1 / c_j, scaled exactly by 2^-7)
and, for the simulation kernels' FP64 sine and cosine (not part of the noise stream), the double nearest to the
exact value (Taylor series in 60-digit decimal arithmetic, pi to 63 digits):
  * SP_TRIG64_TAB[k] = (sin(k pi / 64), cos(k pi / 64), and the doubles nearest to the remainders sin - hi and
    cos - hi), k = 0..127
  * SP_ATAN32_TAB[k] = (atan(k / 32), the double nearest to the remainder atan - hi), k = 0..32
"""
import math
from decimal import Decimal, getcontext
from pathlib import Path

getcontext().prec = 60
PI_DEC = Decimal("3.14159265358979323846264338327950288419716939937510582097494459")


def dec_sincos(x):
    """sin x and cos x to about 58 digits (Taylor series, |x| < 2 pi)."""
    s = c = Decimal(0)
    t, n = Decimal(1), 0          # t = x^n / n!
    while n < 200:
        if n % 4 == 0:
            c += t
        elif n % 4 == 1:
            s += t
        elif n % 4 == 2:
            c -= t
        else:
            s -= t
        n += 1
        t = t * x / n
    return s, c


def dec_atan(x):
    """atan x to about 58 digits for 0 <= x <= 1: three argument halvings atan x = 2 atan(x / (1 + sqrt(1 + x^2))),
    then the Taylor series."""
    halvings = 3
    for _ in range(halvings):
        x = x / (1 + (1 + x * x).sqrt())
    s, t, n = Decimal(0), x, 0
    while n < 400:
        s += t / (2 * n + 1) if n % 2 == 0 else -t / (2 * n + 1)
        t = t * x * x
        n += 1
    return s * (2 ** halvings)


def nearest(d):
    """The double nearest to d; the zeros of sin and cos (left as ~1e-60 by the 63-digit pi) are exact zeros."""
    if abs(d) < Decimal("1e-40"):
        return 0.0
    return float(format(d, ".45e"))          # CPython rounds decimal strings correctly

TWO_PI = 6.283185307179586  # nearest double to 2 pi


def hexd(x: float) -> str:
    return float(x).hex()


def main():
    lines = ["// sp_rng_tables.h -- generated by gen_rng_tables.py; do not edit.",
             "// Constant tables of the FP64 Box-Muller transform of sp_rng.h.",
             "#pragma once", "", "namespace sp {", "",
             "// (sin t_j, cos t_j), t_j = (j + 1/2) (2 pi / 128)",
             "static constexpr double SP_SINCOS_TAB[128][2] = {"]
    step = TWO_PI / 128.0                       # exact: division by a power of two
    for j in range(128):
        t = (j + 0.5) * step
        lines.append(f"    {{{hexd(math.sin(t))}, {hexd(math.cos(t))}}},")
    lines += ["};", "", "// (log c_j, 2^-7 / c_j), c_j = 1 + j / 128, row j + 32 for j = -32..64",
              "static constexpr double SP_LOG_TAB[97][2] = {"]
    for j in range(-32, 65):
        c = 1.0 + j / 128.0                     # exact
        lines.append(f"    {{{hexd(math.log(c))}, {hexd((1.0 / c) * 2.0 ** -7)}}},")
    lines += ["};", "", "// (sin_hi, cos_hi, sin_lo, cos_lo)(k pi / 64), k = 0..127: hi the nearest double to the exact value,",
              "// lo the nearest double to the exact value - hi",
              "static constexpr double SP_TRIG64_TAB[128][4] = {"]
    for k in range(128):
        sk, ck = dec_sincos(PI_DEC * k / 64)
        sh, ch = nearest(sk), nearest(ck)
        sl, cl = nearest(sk - Decimal(sh)), nearest(ck - Decimal(ch))
        lines.append(f"    {{{hexd(sh)}, {hexd(ch)}, {hexd(sl)}, {hexd(cl)}}},")
    lines += ["};", "", "// (atan_hi, atan_lo)(k / 32), k = 0..32", "static constexpr double SP_ATAN32_TAB[33][2] = {"]
    for k in range(33):
        a = dec_atan(Decimal(k) / 32)
        hi = nearest(a)
        lines.append(f"    {{{hexd(hi)}, {hexd(nearest(a - Decimal(hi)))}}},")
    lines += ["};", "", "}  // namespace sp", ""]
    out = Path(__file__).resolve().parent / "sp_rng_tables.h"
    out.write_text("\n".join(lines))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
