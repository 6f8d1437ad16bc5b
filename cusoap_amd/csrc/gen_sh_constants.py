"""Writes sh_constants.inc: normalisation constants of the real solid harmonics.

K[l][m] = sqrt((2l+1)/(4 pi) (l-m)!/(l+m)!) times sqrt(2) for m > 0, for
l <= LCAP; plus the double factorials (2m-1)!! used to seed the recurrence.
Run from this folder: python gen_sh_constants.py
"""
import math

LCAP = 12
lines = ["// generated by gen_sh_constants.py -- do not edit", f"#define CUSOAP_SH_LCAP {LCAP}",
         "static constexpr double kSHNorm[CUSOAP_SH_LCAP + 1][CUSOAP_SH_LCAP + 1] = {"]
for l in range(LCAP + 1):
    row = []
    for m in range(LCAP + 1):
        if m > l:
            row.append("0.0")
            continue
        k = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m))
        if m > 0:
            k *= math.sqrt(2.0)
        row.append(repr(k))
    lines.append("    {" + ", ".join(row) + "},")
lines.append("};")
df = []
for m in range(LCAP + 1):
    v = 1.0
    for k in range(1, 2 * m, 2):
        v *= k
    df.append(repr(v))
lines.append("static constexpr double kDoubleFact[CUSOAP_SH_LCAP + 1] = {" + ", ".join(df) + "};")
with open("sh_constants.inc", "w") as fh:
    fh.write("\n".join(lines) + "\n")
