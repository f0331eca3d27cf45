"""The C-ABI library loads and exports every symbol include/bbwadg.h declares (no GPU calls)."""
import ctypes
import os
import re

from bbwadg_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            text = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"^\s*(?:int|void|double|const char \*)\s*\*?(bbw_\w+)\s*\(", text, re.M))
    return names


def test_header_declares_exports():
    assert declared_symbols() == set(_lib.EXPORTS)


def test_library_exports_all_symbols():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_bad_config_rejected_without_gpu_work():
    lib = _lib.load()
    h = ctypes.c_void_p()
    cfg = _lib.Config(N=10, Nfields=4, K=10, Nq=95)
    assert lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BBW_ERR_UNSUPPORTED
    cfg = _lib.Config(N=7, Nfields=9, K=10, Nq=214)   # elastic kernels: N <= 6
    assert lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BBW_ERR_UNSUPPORTED
    cfg = _lib.Config(N=9, Nfields=4, K=10, Nq=729)   # N = 9 has 1000 WADG points
    assert lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BBW_ERR_ARG
    cfg = _lib.Config(N=5, Nfields=5, K=10, Nq=95)
    assert lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BBW_ERR_UNSUPPORTED
    cfg = _lib.Config(N=5, Nfields=4, K=10, Nq=94)
    assert lib.bbw_create(ctypes.byref(cfg), ctypes.byref(h)) == _lib.BBW_ERR_ARG
    assert lib.bbw_create(None, ctypes.byref(h)) == _lib.BBW_ERR_ARG


def test_header_constants_match_bindings():
    """Every option, mode and implementation constant of the header has the same value in the
    Python bindings."""
    text = open(os.path.join(ROOT, "include", "bbwadg.h")).read()
    defs = dict((k, int(v)) for k, v in re.findall(r"^#define BBW_(\w+) \(?(-?\d+)\)?", text, re.M))
    checked = 0
    for k, v in defs.items():
        if hasattr(_lib, k):
            assert getattr(_lib, k) == v, k
            checked += 1
        elif hasattr(_lib, "BBW_" + k):
            assert getattr(_lib, "BBW_" + k) == v, k
            checked += 1
    assert checked >= 12
