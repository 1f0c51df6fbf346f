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
    assert lib.ar_abi_version() == 1


def test_null_arguments_rejected():
    lib = _lib.load()
    assert lib.ar_create(None, None) == _lib.AR_E_INVALID
    assert lib.ar_destroy(None) == _lib.AR_E_INVALID
    assert lib.ar_scene_set(None, None, 0) == _lib.AR_E_INVALID
    assert lib.ar_bvh_build(None, None) == _lib.AR_E_INVALID
    assert lib.ar_trace_closest(None, None, 0, None, None) == _lib.AR_E_INVALID
    assert lib.ar_last_error(None) == b"null context"
