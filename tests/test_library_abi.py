"""CPU checks of the C ABI: the library loads and exports every entry point
declared in include/ccsd_gpu_amd.h (no compute calls without a GPU)."""
import ctypes
import os
import re

from ccsd_gpu_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "ccsd_gpu_amd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ccsd_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_loader_signatures():
    assert set(declared_symbols()) == set(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_errors_without_device_are_reported_not_crashing():
    lib = _lib.load()
    h = ctypes.c_void_p()
    assert lib.ccsd_plan_create(ctypes.byref(h), None, 0, None, 0, None, 0, None) == -1
    assert b"null" in lib.ccsd_last_error()
    assert lib.ccsd_plan_destroy(None) == 0
    assert lib.ccsd_rys_roots(10, None, 0, None, None, None) == -1
