"""The C ABI library loads and exports every symbol include/method_gpu.h declares,
and argument validation fails with the reference's messages before any device
call (so these run without a GPU)."""
import ctypes
import os
import re

import pytest

from method_amd import engine as eng

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "method_gpu.h")).read()
    return sorted(set(re.findall(r"\b(mgpu_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_python_binding_list():
    assert declared_symbols() == sorted(eng.EXPORTED)


def test_library_exports_every_declared_symbol():
    L = eng.load_library()
    for name in declared_symbols():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def _make(**kw):
    base = dict(nx=16, ny=16, nz=0, xmin=0.0, xmax=1.0, ymin=0.0, ymax=1.0)
    base.update(kw)
    return eng.Config(**base)


@pytest.mark.parametrize("kw, msg", [
    (dict(weno_order=5, Ng=3), "You must increase number of boundary cells."),  # weno.cc:13-20
    (dict(sigma=-1.0), "Conductivity must be non-negative"),                   # simData.cc:100-102
    (dict(weno_order=7), "weno_order must be 3 or 5"),
    (dict(nx=0), "need nx >= 1, ny >= 0, nz >= 0"),
    (dict(ny=0, nz=8), "a grid with nz > 0 needs ny > 0"),
])
def test_create_rejects_bad_configuration(kw, msg):
    with pytest.raises(eng.EngineError, match=re.escape(msg)):
        eng.Engine(_make(**kw))


def test_missing_library_raises(tmp_path):
    saved = eng._lib
    try:
        eng._lib = None
        with pytest.raises(eng.EngineError, match="not built"):
            eng.load_library(str(tmp_path / "nope.so"))
    finally:
        eng._lib = saved
