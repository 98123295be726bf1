"""The C-ABI library loads and exports every entry point include/nsim.h declares
(no compute: these run without a GPU)."""
import ctypes
import os
import re

from nmag import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "nsim.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(nsim_[a-z_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    names = declared_symbols()
    assert len(names) >= 15
    for name in names:
        assert hasattr(lib, name), name


def test_python_binding_covers_the_header():
    assert set(declared_symbols()) <= set(_native.SIGNATURES)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setattr(_native, "LIB_PATH", "/nonexistent/libnsim_hip.so")
    try:
        _native.lib()
    except ImportError as e:
        assert "no fallback" in str(e)
    else:
        raise AssertionError("expected ImportError")
