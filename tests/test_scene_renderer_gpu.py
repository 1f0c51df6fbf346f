"""The SceneRenderer / Camera front end on the GPU: the C++ program of
test_scene_renderer.py (tests/cxx/scene_renderer_main.cpp) linked against
libar_hip.so, its interactive buffers in device memory, checked against the
FP32 oracle (bars in tests/scene_program.py: the integrator's relMSE and
per-pixel bars, display images within one level)."""
import pytest

import scene_program as sp

pytestmark = pytest.mark.gpu


def test_scene_renderer_program_on_gpu(tmp_path):
    out = sp.run("hip", tmp_path, timeout=300)
    sp.check(out, exact=False)
