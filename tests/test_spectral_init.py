"""Initial parameters of the spectral basis network (host code, no GPU):
the reference trainer's initialisation (train_spectrum.py:103-160) laid out
in the parameter order the library uses."""
import numpy as np

import oracle_lib as orc
from aperturerendererv3_amd import _lib, spectral_mlp_init


def test_layout_and_statistics():
    w = spectral_mlp_init(3)
    assert w.dtype == np.float32 and w.size == _lib.SMLP_PARAMS == orc.SMLP_PARAMS
    decays = orc.smlp_decay_mask()
    assert np.all(w[~decays] == 0.0)                       # every bias starts at zero
    mats = [w[0:1024]] + [w[1056 + k * 1056:1056 + k * 1056 + 1024] for k in range(7)] + \
           [w[8576 + k * 1056:8576 + k * 1056 + 1024] for k in range(7)]
    allm = np.concatenate(mats)
    assert abs(allm.mean()) < 0.01 and abs(allm.std() - np.sqrt(2.0 / 32.0)) < 0.01   # Kaiming normal, fan in 32
    for out in (w[8448:8576], w[15968:16096]):             # torch's default Linear init
        assert np.abs(out).max() <= 1.0 / np.sqrt(32.0) and out.std() > 0.07   # uniform: std 0.102
    assert np.array_equal(spectral_mlp_init(3), w) and not np.array_equal(spectral_mlp_init(4), w)
