"""Writes t_quantile_table.inc: the 1 - 0.005 / 2 quantile of Student's t
distribution for nu = 1 .. 1024 degrees of freedom (two-sided 0.5 % level),
as float literals, for the denoiser's Welch t-test (the role of
T005QuantilesDev, core/denoise/t_quantile.cu).  Computed with scipy; run once
when the table has to change:  python aperturerendererv3_amd/csrc/gen_t_quantiles.py"""
from pathlib import Path

import numpy as np
from scipy.stats import t

OUT = Path(__file__).resolve().parent / "t_quantile_table.inc"


def main():
    nu = np.arange(1, 1025, dtype=np.float64)
    q = t.ppf(1.0 - 0.005 / 2.0, nu).astype(np.float32)
    lines = ["// Student t quantile at 1 - 0.005 / 2 for nu = 1 .. 1024 (scipy.stats.t.ppf), generated by",
             "// gen_t_quantiles.py; entry nu - 1 is nu degrees of freedom."]
    for k in range(0, 1024, 6):
        lines.append(" ".join(f"{float(v)!r}f," for v in q[k:k + 6]))
    OUT.write_text("\n".join(lines) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
