"""Generate sp_rng_tables.h: the constant tables of the FP64 Box-Muller transform.

    python stochasticphysics_amd/csrc/gen_rng_tables.py     # rewrites sp_rng_tables.h beside this file

The tables are part of the noise-stream definition (sp_rng.h), so the header
is committed and the build never regenerates it.  Every entry is the double
nearest to the stated value as Python's math module (the C library)
computes it from the given double arguments:
  * SP_SINCOS_TAB[j] = (sin t_j, cos t_j), t_j = (j + 1/2) * (2 pi / 128), j = 0..127
  * SP_LOG_TAB[j + 32] = (log c_j, 2^-7 / c_j), c_j = 1 + j / 128, j = -32..64
    (2^-7 / c_j is the double nearest to 1 / c_j, scaled exactly by 2^-7)
"""
import math
from pathlib import Path

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
    lines += ["};", "", "}  // namespace sp", ""]
    out = Path(__file__).resolve().parent / "sp_rng_tables.h"
    out.write_text("\n".join(lines))
    print(f"wrote {out}")


if __name__ == "__main__":
    main()
