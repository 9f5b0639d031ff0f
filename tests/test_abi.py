"""The C ABI library loads and exports every symbol include/m70h.h declares (no GPU
calls), and error paths that need no device."""
import ctypes
import os
import re

from morfo70_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "m70h.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(m70[hswo]?(?:nc)?_[a-z_0-9]+)\s*\(", txt)))


def test_header_matches_binding():
    assert declared_symbols() == sorted(_lib.SIGNATURES)


def test_library_exports_all_symbols():
    L = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(L, name), name


def test_create_rejects_bad_dims_without_device():
    L = _lib.load()
    h = ctypes.c_void_p()
    x = (ctypes.c_double * 2)(0.0, 1.0)
    p = _lib.Params(0, 0, 0, 0.5, 0.01, 1.16, 0.002, 1e-4, 1.0, 0.02, 0.001, 1.0)
    rc = L.m70h_create(2, 8, ctypes.cast(x, ctypes.POINTER(ctypes.c_double)), 1.0,
                       ctypes.cast(ctypes.pointer(p), ctypes.c_void_p), ctypes.byref(h))
    assert rc == _lib.M70H_EINVAL
    assert b"nx and ny" in L.m70h_last_error(None)
    assert not h.value


def test_sediment_create_rejects_null_hydro_without_device():
    L = _lib.load()
    s = ctypes.c_void_p()
    rc = L.m70s_create(None, None, None, None, ctypes.byref(s))
    assert rc == _lib.M70H_EINVAL
    assert b"non-NULL" in L.m70s_last_error(None)
    assert not s.value
