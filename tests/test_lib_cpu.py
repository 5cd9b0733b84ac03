"""The C-ABI library without a GPU: it loads, exports every function that
include/dg.h declares, and rejects bad configurations before touching HIP."""
import ctypes
import os
import re

import pytest

from dgcuda_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "dg.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(dg_\w+)\s*\(", txt))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(L.LIB_PATH):
        from dgcuda_amd import build
        build.build()
    return L.load()


def test_exports_every_declared_function(lib):
    declared = header_functions()
    assert declared == set(L.EXPORTS)
    raw = ctypes.CDLL(L.LIB_PATH)
    for name in declared:
        assert hasattr(raw, name), name


def test_config_defaults_and_validation(lib):
    cfg = L.DgConfig()
    lib.dg_config_default(ctypes.byref(cfg))
    assert cfg.physics == L.EULER and cfg.integrator == L.RK4 and cfg.gamma == 1.4
    assert cfg.g == 9.8  # shallowwater.cu:7
    h = ctypes.c_void_p()
    for field, bad in (("p", 6), ("physics", 9), ("integrator", 0)):
        c = L.DgConfig()
        lib.dg_config_default(ctypes.byref(c))
        setattr(c, field, bad)
        assert lib.dg_create(ctypes.byref(h), ctypes.byref(c)) == -1
    c = L.DgConfig()
    lib.dg_config_default(ctypes.byref(c))
    c.limiter, c.p = 1, 2  # the limiter is p = 1 only (main.cu:647-650)
    assert lib.dg_create(ctypes.byref(h), ctypes.byref(c)) == -1
    c.limiter, c.p, c.flux = 0, 2, L.FLUX_ROE  # Roe is shallow water only
    assert lib.dg_create(ctypes.byref(h), ctypes.byref(c)) == -1
    # the matrix-core volume kernel exists for Euler p = 4 only
    for physics, p, vol in ((L.EULER, 3, L.VOLUME_MFMA), (L.SWE, 4, L.VOLUME_MFMA),
                            (L.EULER, 4, 3), (L.EULER, 4, -1)):
        c = L.DgConfig()
        lib.dg_config_default(ctypes.byref(c))
        c.physics, c.p, c.volume = physics, p, vol
        assert lib.dg_create(ctypes.byref(h), ctypes.byref(c)) == -1


def test_null_handle_calls_are_rejected(lib):
    assert lib.dg_project_ic(None) == -1
    assert lib.dg_step(None, 1, 0.0, None, None) == -1
    assert lib.dg_last_error(None) == b"null handle"


def test_missing_library_fails_loudly(monkeypatch):
    """no fallback: without libdgamd.so the binding and the solver raise"""
    from dgcuda_amd import _lib as Lb
    from dgcuda_amd.solver import Solver
    monkeypatch.setattr(Lb, "_lib", None)
    monkeypatch.setattr(Lb, "LIB_PATH", "/nonexistent/libdgamd.so")
    with pytest.raises(OSError):
        Lb.load()
    with pytest.raises(OSError):
        Solver(Lb.EULER, 2)
