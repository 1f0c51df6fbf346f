"""Writes t005_quantiles_reference.npy: the reference's table of Student t
quantiles used by its denoiser (T005QuantilesHost, core/denoise/t_quantile.cu,
1024 float32 values, entry nu - 1 for nu degrees of freedom), read once from
the reference checkout.  The library's own table is computed with scipy
(aperturerendererv3_amd/csrc/gen_t_quantiles.py); tests compare the two."""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/core/denoise/t_quantile.cu")
OUT = Path(__file__).resolve().parent / "t005_quantiles_reference.npy"


def main():
    text = SRC.read_text()
    block = text[text.index("#define T005_LUT_INIT {") + len("#define T005_LUT_INIT {"):]
    block = block[:block.index("}")]
    vals = np.array([float(v.strip().rstrip("f")) for v in block.split(",") if v.strip()], np.float32)
    assert vals.shape == (1024,), vals.shape
    np.save(OUT, vals)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
