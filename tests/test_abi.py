"""The C ABI: liborbit_mi355x.so loads without a GPU and exports every function declared
in include/orbit_mi355x.h (no compute calls here)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    txt = open(os.path.join(ROOT, "include", "orbit_mi355x.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(?:int|const char\*)\s+(ob_\w+)\s*\(", txt)))


def test_library_exports_declared_symbols():
    lib = ctypes.CDLL(os.path.join(ROOT, "py-orbit_amd", "liborbit_mi355x.so"))
    names = declared()
    assert len(names) > 50
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_bindings_cover_the_abi():
    import _obnative
    names = set(declared())
    bound = set(_obnative.SIGNATURES) | set(_obnative._RET)
    assert names <= bound, sorted(names - bound)


def test_version_and_error_without_device():
    import _obnative
    L = _obnative.load()
    assert b"gfx950" in L.ob_version()
    assert isinstance(L.ob_last_error(), bytes)
