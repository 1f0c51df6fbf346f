"""Writes spectral_mlp_reference.npz from the reference checkout
(core/spectrum/spectrum_basis.cuh): the shipped weights of its spectral basis
network and its 100 golden vectors, used only as test inputs and expected
outputs for the network's forward pass.

  weights   16096 float32 in the layer order of the header
            (RGB_IN_W, RGB_IN_B, RGB_BLK0_W, RGB_BLK0_B, ..., RGB_BLK6_B,
            RGB_OUT_W, LAM_BLK0_W, LAM_BLK0_B, ..., LAM_BLK6_B, LAM_OUT_W),
            each weight matrix row-major [out][in]
  inputs    [100, 4]: r, g, b, wavelength in nm (inputTest, transposed)
  expected  [100]: gtTest, the reflectance the trained network returned for
            each input (train_spectrum.py:417-443 evaluates the model on them)
"""
from pathlib import Path

import numpy as np

SRC = Path("/root/reference/core/spectrum/spectrum_basis.cuh")
OUT = Path(__file__).resolve().parent / "spectral_mlp_reference.npz"

LAYERS = (["RGB_IN_W", "RGB_IN_B"] + [f"RGB_BLK{k}_{p}" for k in range(7) for p in "WB"] + ["RGB_OUT_W"]
          + [f"LAM_BLK{k}_{p}" for k in range(7) for p in "WB"] + ["LAM_OUT_W"])


def _body(text, name):
    head = f"inline constexpr float {name}["
    at = text.index(head)
    start = text.index("{", text.index("#endif", at))
    depth, i = 0, start
    while True:
        c = text[i]
        depth += c == "{"
        depth -= c == "}"
        if depth == 0:
            return text[start + 1:i]
        i += 1


def _floats(s):
    return np.array([float(v.strip().rstrip("f")) for v in s.replace("{", ",").replace("}", ",").split(",")
                     if v.strip()], np.float32)


def main():
    text = SRC.read_text()
    w = np.concatenate([_floats(_body(text, n)) for n in LAYERS])
    assert w.size == 16096, w.size
    inputs = _floats(_body(text, "inputTest")).reshape(4, 100).T.copy()
    expected = _floats(_body(text, "gtTest"))
    assert expected.shape == (100,)
    np.savez(OUT, weights=w, inputs=inputs, expected=expected)
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
