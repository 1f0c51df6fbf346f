"""The SceneRenderer / Camera front end (include/ar/scene_renderer.hpp)
through a C++ program (tests/cxx/scene_renderer_main.cpp) linked against
the oracle's libar_cpu.so: offline render, RenderToFile, interactive frames
with odd / even halves, denoiser records, the denoised and albedo displays, a
voxel-guided render, SDF volumes from a host functor and from the library's
bake, an error path and camera motion.  The CPU backend is the oracle, so
every image equals the oracle's bit for bit; the same program and checks run
against libar_hip.so in test_scene_renderer_gpu.py."""
import subprocess

import scene_program as sp


def test_scene_renderer_program_on_cpu_backend(tmp_path):
    subprocess.run(["make", "-s", "-C", str(sp.ROOT / "oracle"), "libar_cpu.so"], check=True)
    out = sp.run("cpu", tmp_path)
    sp.check(out, exact=True)


def test_wrapper_compiles_as_cxx14(tmp_path):
    """The header needs C++14 only and no HIP header."""
    src = tmp_path / "t.cpp"
    src.write_text('#include "ar/scene_renderer.hpp"\nint main() { SceneObject o; return (int)o.GetArea(); }\n')
    subprocess.run(["g++", "-std=c++14", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", f"-I{sp.ROOT / 'include'}",
                    str(src)], check=True)
