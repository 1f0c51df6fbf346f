"""Writes the CIE 1931 2-degree colour-matching functions and the CIE D65
spectral power distribution, 360-780 nm at 1 nm, to
aperturerendererv3_amd/data/cie1931_d65_1nm.txt.

These are public CIE standard tables (CIE 1931 standard colorimetric
observer; CIE standard illuminant D65).  The reference renderer tabulates the
same standard values (core/spectrum/ciexyz.cuh:12-95, 96-520); this script
reads the numbers from there once, in the build container, because the image
has no other copy of the standard.  Output columns: lambda_nm xbar ybar zbar
d65.
"""
import re
import sys
from pathlib import Path

SRC = Path("/root/reference/core/spectrum/ciexyz.cuh")
OUT = Path(__file__).resolve().parents[2] / "aperturerendererv3_amd" / "data" / "cie1931_d65_1nm.txt"


def main():
    text = SRC.read_text()
    d65_block = text[text.index("D65_SPD[D65_COUNT] =\n#endif"):]
    d65_block = d65_block[d65_block.index("{") + 1:d65_block.index("};")]
    d65 = [float(x.rstrip("f")) for x in re.findall(r"[-0-9.eE+]+f", d65_block)]
    cmf_block = text[text.index("CIEXYZ_CMF[CIEXYZ_COUNT] =\n#endif"):]
    cmf_block = cmf_block[cmf_block.index("{") + 1:cmf_block.index("};")]
    triples = re.findall(r"float3\{([^}]*)\}", cmf_block)
    cmf = [[float(v.strip().rstrip("f")) for v in t.split(",")] for t in triples]
    assert len(d65) == 421 and len(cmf) == 421, (len(d65), len(cmf))
    lines = ["# CIE 1931 2-degree observer (xbar ybar zbar) and CIE illuminant D65, 360-780 nm, 1 nm",
             "# lambda_nm xbar ybar zbar d65"]
    for i in range(421):
        x, y, z = cmf[i]
        lines.append(f"{360 + i} {x!r} {y!r} {z!r} {d65[i]!r}")
    OUT.write_text("\n".join(lines) + "\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    sys.exit(main())
