"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (the GPU
code cannot be: GPU ASan and XNACK are not available on the MI355X pool).

- The BVH builder and the 4- / 8-wide collapse of the HIP library
  (bvh_build.cpp, bvh_wide.cpp, compiled by g++ with the sanitizers) on the
  meshes and settings of tests/cxx/bvh_host_check.cpp, which also checks the
  trees' structural invariants.
- The oracle (oracle/liboracle_san.so, loaded with libasan / libubsan
  preloaded into the test process) under its own CPU tests and the
  fault-injection cases.
- The SceneRenderer / Camera wrapper and the C-ABI CPU backend
  (libar_cpu_san.so) through the wrapper program, with the bit-exact checks
  of test_scene_renderer.py.
Any sanitizer report fails the run (UBSan without recovery, ASan aborts)."""
import os
import subprocess
import sys
from pathlib import Path

import scene_program as sp

ROOT = Path(__file__).resolve().parent.parent
BUILD = ROOT / "tests" / "cxx" / "build"
SAN = ["-g", "-O1", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
REPORTS = ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer")


def _gcc_lib(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True, check=True).stdout
    return path.strip()


def _assert_clean(p):
    text = p.stdout + p.stderr
    print(text[-3000:], flush=True)
    assert p.returncode == 0, text[-3000:]
    for r in REPORTS:
        assert r not in text, r


def test_host_builder_under_sanitizers():
    BUILD.mkdir(parents=True, exist_ok=True)
    exe = BUILD / "bvh_host_check_san"
    csrc = ROOT / "aperturerendererv3_amd" / "csrc"
    subprocess.run(["g++", "-std=c++17", *SAN, "-D__HIP_PLATFORM_AMD__", f"-I{ROCM}/include",
                    str(ROOT / "tests" / "cxx" / "bvh_host_check.cpp"), str(csrc / "bvh_build.cpp"),
                    str(csrc / "bvh_wide.cpp"), "-o", str(exe), "-lpthread"], check=True)
    p = subprocess.run([str(exe)], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    _assert_clean(p)
    assert p.stdout.strip().endswith("ok")


ORACLE_TESTS = ["test_oracle_trace.py", "test_oracle_bsdf.py", "test_oracle_render.py", "test_oracle_objects.py",
                "test_oracle_render_objects.py", "test_oracle_output.py", "test_oracle_denoise.py",
                "test_oracle_guiding.py", "test_oracle_mlp.py", "test_oracle_sdf.py", "test_oracle_spectral.py",
                "test_oracle_spectral_mlp.py", "test_fault_injection.py"]


def test_oracle_under_sanitizers():
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "liboracle_san.so"], check=True)
    preload = [_gcc_lib("libasan.so"), _gcc_lib("libubsan.so")]
    if os.environ.get("LD_PRELOAD"):
        preload.append(os.environ["LD_PRELOAD"])
    env = dict(os.environ, LD_PRELOAD=":".join(preload), AR_ORACLE_LIB="liboracle_san.so",
               ASAN_OPTIONS="detect_leaks=0", UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    tests = [str(ROOT / "tests" / t) for t in ORACLE_TESTS if (ROOT / "tests" / t).exists()]
    # the slowest statistical cases (unbiasedness of guided renders, the
    # descent fit) run without the sanitizers in the ordinary suite
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        "-k", "not cpu_backend and not unbiased and not descent", *tests],
                       capture_output=True, text=True, timeout=1200, env=env, cwd=str(ROOT))
    _assert_clean(p)
    assert " passed" in p.stdout


def test_wrapper_and_cpu_backend_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "libar_cpu_san.so"], check=True)
    out = sp.run("cpu-san", tmp_path)   # a sanitizer report exits non-zero (no recovery)
    sp.check(out, exact=True)
