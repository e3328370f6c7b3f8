"""The C ABI library: loads without a GPU, exports every entry point that
include/axr.h declares, and the product path refuses to run without it."""
import ctypes
import os
import re

import pytest

from axi_visualizer_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "axr.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(axr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_bound_functions():
    names = declared_symbols()
    assert len(names) >= 25
    assert set(names) == set(_native.SIGNATURES), \
        set(names) ^ set(_native.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_bad_arguments_fail_without_touching_the_gpu():
    L = _native.lib()
    out = ctypes.c_void_p()
    assert L.axr_create(0, 64, ctypes.byref(out)) == -1
    assert b"must be positive" in L.axr_last_error()
    assert L.axr_set_scale_cull(None, 1.0, 1.0, 1.0) == -1
    assert L.axr_draw_all(None, None, None, 0, None, 0, 5) == -1


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_native, "LIB_PATH", "/nonexistent/libaxr.so")
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(RuntimeError, match="not built"):
        _native.lib()


def test_shader_strings_map_to_flags():
    from axi_visualizer_amd import shader_flags
    from axi_visualizer_amd.scenes import SH_CULL, SH_MIP, SH_PHONG, SH_SHADOW

    assert shader_flags("mip shadow  ") == SH_MIP | SH_SHADOW
    assert shader_flags("phong cull") == SH_PHONG | SH_CULL
    assert shader_flags({"mip": 0.5}) == SH_MIP
    assert shader_flags(5) == 5
    for kw in ("alpha", "sky", "SSR", "refl_x", "multi"):
        with pytest.raises(NotImplementedError):
            shader_flags("mip " + kw)
