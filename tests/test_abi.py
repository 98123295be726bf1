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


def test_batch_create_rejects_invalid_meshes_before_touching_the_device():
    """Input validation of nsim_create_batch runs on the host (no GPU needed)."""
    import numpy as np
    from nmag import kuhn_cuboid
    L = _native.lib()
    c0, t0 = kuhn_cuboid(2, 1, 1, 2.5)
    c1, t1 = kuhn_cuboid(1, 1, 1, 2.5)
    t1 = t1.copy()
    t1[0, 0] = len(c1)  # a node of the next magnet: must be rejected
    keep = []
    meshes = (_native.NsimMesh * 2)()
    mats = (_native.NsimMaterial * 2)()
    for k, (c, t) in enumerate([(c0, t0), (c1, t1)]):
        c = np.ascontiguousarray(c, dtype=np.float64)
        t = np.ascontiguousarray(t, dtype=np.int32)
        keep += [c, t]
        meshes[k] = _native.NsimMesh(len(c), len(t), _native._dptr(c),
                                     t.ctypes.data_as(_native._ip))
        mats[k] = _native.NsimMaterial(0.86, -24.0, -0.2, -0.1, 0.1, 1.2566)
    h = ctypes.c_void_p()
    rc = L.nsim_create_batch(2, meshes, mats, None, ctypes.byref(h))
    assert rc < 0 and not h.value
    assert b"magnet 1" in L.nsim_last_error() and b"out of range" in L.nsim_last_error()
    assert L.nsim_create_batch(0, meshes, mats, None, ctypes.byref(h)) < 0


def test_create_rejects_unknown_precision_before_touching_the_device():
    """The precision code is checked on the host: 0-3 are defined (FP64,
    FP32-mixed, BF16, 24-bit potential operator), anything else is an error."""
    import numpy as np
    from nmag import kuhn_cuboid
    L = _native.lib()
    c, t = kuhn_cuboid(1, 1, 1, 2.5)
    c = np.ascontiguousarray(c, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.int32)
    mesh = _native.NsimMesh(len(c), len(t), _native._dptr(c), t.ctypes.data_as(_native._ip))
    mat = _native.NsimMaterial(0.86, -24.0, -0.2, -0.1, 0.1, 1.2566)
    assert _native.PRECISIONS["mixed24"] == 3
    for bad in (-1, 4):
        opt = _native.NsimOptions(bad, 1, 1e-6, 1e-6, 100, 100, 0)
        h = ctypes.c_void_p()
        rc = L.nsim_create(ctypes.byref(mesh), ctypes.byref(mat), ctypes.byref(opt),
                           ctypes.byref(h))
        assert rc < 0 and not h.value
        assert b"unknown precision" in L.nsim_last_error()
