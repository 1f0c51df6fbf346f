"""The C ABI library loads and exports every entry point include/ar.h declares
(no compute calls: this runs without a GPU)."""
import ctypes as C

from aperturerendererv3_amd import _lib


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(str(_lib.LIB_PATH))
    declared = _lib.declared_symbols()
    assert len(declared) >= 10
    missing = [s for s in declared if not hasattr(lib, s)]
    assert not missing, f"declared in ar.h but not exported: {missing}"
    assert set(declared) == set(_lib.SIGNATURES), "ctypes table out of sync with ar.h"


def test_abi_version():
    lib = _lib.load()
    assert lib.ar_abi_version() == 2


def test_null_arguments_rejected():
    lib = _lib.load()
    assert lib.ar_create(None, None) == _lib.AR_E_INVALID
    assert lib.ar_destroy(None) == _lib.AR_E_INVALID
    assert lib.ar_scene_set(None, None, 0) == _lib.AR_E_INVALID
    assert lib.ar_bvh_build(None, None) == _lib.AR_E_INVALID
    assert lib.ar_trace_closest(None, None, 0, None, None) == _lib.AR_E_INVALID
    assert lib.ar_last_error(None) == b"null context"


def test_ctypes_structs_match_the_header(tmp_path):
    """Every ctypes mirror has the size and field offsets gcc gives the
    struct of include/ar.h."""
    import pathlib
    import shutil
    import subprocess
    import pytest
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    structs = [n for n in dir(_lib) if n.startswith("ar_") and isinstance(getattr(_lib, n), type)
               and issubclass(getattr(_lib, n), C.Structure)]
    assert {"ar_object", "ar_material", "ar_rays", "ar_render_opts"} <= set(structs)
    inc = pathlib.Path(_lib.__file__).resolve().parent.parent / "include"
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ar.h"', "int main(void) {"]
    for n in structs:
        lines.append(f'printf("{n} size %zu\\n", sizeof({n}));')
        for f, _ in getattr(_lib, n)._fields_:
            lines.append(f'printf("{n} {f} %zu\\n", offsetof({n}, {f}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", str(inc), str(src), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.splitlines():
        n, f, v = line.split()
        got[(n, f)] = int(v)
    for n in structs:
        cls = getattr(_lib, n)
        assert got[(n, "size")] == C.sizeof(cls), n
        for f, _ in cls._fields_:
            assert got[(n, f)] == getattr(cls, f).offset, (n, f)


def test_cpu_backend_exports_every_declared_symbol():
    """oracle/libar_cpu.so implements the same ABI (the tests link programs
    against either library)."""
    import pathlib
    import subprocess
    oracle = pathlib.Path(_lib.__file__).resolve().parent.parent / "oracle"
    subprocess.run(["make", "-s", "-C", str(oracle), "libar_cpu.so"], check=True)
    lib = C.CDLL(str(oracle / "libar_cpu.so"))
    missing = [s for s in _lib.declared_symbols() if not hasattr(lib, s)]
    assert not missing, f"declared in ar.h but not in libar_cpu.so: {missing}"
    assert lib.ar_abi_version() == 2
    assert lib.ar_create(None, None) == _lib.AR_E_INVALID


# Names the INTEGRATION.md snippets take from the caller's code.
SNIPPET_PRELUDE = r"""
#include <cstdio>
#include <vector>
#include "ar.h"
#include "ar/scene_renderer.hpp"
extern ar_ctx* ctx;
extern void* stream;
extern const float* verts;
extern const uint32_t* idx;
extern uint32_t nv, ntri;
extern size_t n;
extern const float *d_ox, *d_oy, *d_oz, *d_tmin, *d_dx, *d_dy, *d_dz, *d_tmax;
extern uint32_t *d_prim, *d_flags, *d_occluded_bits, *d_ldr, *h_ldr;
extern float *d_t, *d_u, *d_v, *d_out_rgb, *d_accum, *d_odd, *d_even, *d_x, *d_y, *d_loss_history, *h_hdr;
extern std::vector<SceneObject> sceneObjects;
extern ar_float3 eulerDegrees;
extern ar_camera cam, cams[48];
extern ar_render_opts opts, opts48[48];
extern float* d_outs[48];
extern int w, h;
extern float exposure;
extern uint32_t xyz_input;
extern ar_float3 *d_hdr3, *d_odd3, *d_even3;
extern unsigned* d_ldr_u;
extern float *d_gbuf3, *d_stats3;
extern double dx, dy, ws, ad;
extern SceneObject so;
extern uint32_t f;
std::vector<SceneObject> CreateScene();
"""

# snippet-local names that differ from the prelude's (the snippets use one
# name for buffers of two types in two places)
SNIPPET_RENAMES = {
    "drop-in": [("d_hdr", "d_hdr3"), ("d_odd", "d_odd3"), ("d_even", "d_even3"), ("d_gbuf", "d_gbuf3"),
                ("d_stats", "d_stats3"), ("d_ldr", "d_ldr_u")],
}


def test_integration_snippets_compile(tmp_path):
    """Every C++ block of INTEGRATION.md compiles against include/ar.h and
    include/ar/scene_renderer.hpp (each block as a function body; its
    #include lines at file scope)."""
    import pathlib
    import re
    import shutil
    import subprocess
    import pytest
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    root = pathlib.Path(_lib.__file__).resolve().parent.parent
    text = (root / "INTEGRATION.md").read_text()
    blocks = re.findall(r"```cpp\n(.*?)```", text, flags=re.S)
    assert len(blocks) >= 8
    src = [SNIPPET_PRELUDE]
    for k, b in enumerate(blocks):
        body = [ln for ln in b.splitlines() if not ln.startswith("#include")]
        code = "\n".join(body)
        if "AddVolumeToCacheManager" in code:     # the drop-in block
            for a, r in SNIPPET_RENAMES["drop-in"]:
                code = re.sub(rf"\b{a}\b", r, code)
        code = code.replace("cams, opts, ", "cams, opts48, ").replace("cams, opts,", "cams, opts48,")
        src.append(f"void snippet_{k}() {{\n{code}\n}}\n")
    f = tmp_path / "snippets.cpp"
    f.write_text("\n".join(src))
    p = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wno-unused-variable",
                        "-Wno-unused-but-set-variable", f"-I{root / 'include'}", str(f)],
                       capture_output=True, text=True)
    assert p.returncode == 0, p.stderr[-4000:]
