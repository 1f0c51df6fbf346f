"""Build libar_hip.so (HIP kernels for gfx950 + the C ABI) in-tree.

Usage: python -m aperturerendererv3_amd.build   (or build() from Python)

Every HIP translation unit is compiled for gfx950 only, without FMA
contraction by default (the triangle test must round like the CPU formula;
kernels opt into FMA explicitly with fmaf), and with the packed-FP32 feature
turned off.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OUT = PKG / "libar_hip.so"
BUILD_DIR = PKG / "_build"
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Wall", "-Wno-unused-function",
    f"-I{PKG.parent / 'include'}",
    f"-I{PKG / '_build'}",
]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


def _hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _sources():
    return sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))


# Per-source device flags.  mlp.hip: MFMA accumulators in VGPRs, not AGPRs
# (gfx950 allows either), which saves the guiding network's inference kernel
# its 96 accumulator-to-VGPR copies per 32-vertex tile.
SOURCE_FLAGS = {src: ["-mllvm", "-amdgpu-mfma-vgpr-form"] for src in ("mlp.hip", "smlp.hip")}


def _compile(src: Path) -> Path:
    obj = BUILD_DIR / (src.name + ".o")
    cmd = [_hipcc(), "-c", str(src), "-o", str(obj)] + COMMON + DEVICE
    if src.suffix == ".hip":
        cmd = [_hipcc(), "-x", "hip", "-c", str(src), "-o", str(obj)] + COMMON + DEVICE + SOURCE_FLAGS.get(src.name, [])
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


DATA = PKG / "data" / "cie1931_d65_1nm.txt"


def _gen_tables():
    """Embeds the CIE 1931 / D65 table (data/cie1931_d65_1nm.txt) as C arrays."""
    rows = [ln.split() for ln in DATA.read_text().splitlines() if ln.strip() and not ln.startswith("#")]
    assert len(rows) == 421
    xyz = ",\n".join(f"{float(r[1])!r}f, {float(r[2])!r}f, {float(r[3])!r}f" for r in rows)
    d65 = ",\n".join(f"{float(r[4])!r}f" for r in rows)
    text = ("// generated from data/cie1931_d65_1nm.txt by build.py\n"
            f"static const float kCieXyz1nm[421 * 3] = {{\n{xyz}\n}};\n"
            f"static const float kD65_1nm[421] = {{\n{d65}\n}};\n")
    out = BUILD_DIR / "cie_tables.inc"
    if not out.exists() or out.read_text() != text:
        out.write_text(text)


def build(verbose: bool = False) -> Path:
    BUILD_DIR.mkdir(exist_ok=True)
    _gen_tables()
    srcs = _sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [_hipcc(), "-shared", "-o", str(tmp)] + [str(o) for o in objs] + DEVICE + ["-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
