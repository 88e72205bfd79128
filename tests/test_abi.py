"""The C ABI library loads and exports every symbol include/gpjson.h declares (no GPU calls)."""
import ctypes
import os
import re

import pytest

from gpjson_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "gpjson.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gpjson_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_core_entry_points():
    syms = declared_symbols()
    for s in ["gpjson_ctx_create", "gpjson_index_build", "gpjson_query_run", "gpjson_query_file",
              "gpjson_index_export", "gpjson_last_error", "gpjson_compile"]:
        assert s in syms


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="libgpjson_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_python_binding_covers_header():
    assert set(declared_symbols()) <= set(_native.SIGNATURES), set(declared_symbols()) - set(_native.SIGNATURES)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="libgpjson_hip.so not built")
def test_last_error_is_never_null():
    lib = _native.load()
    assert isinstance(lib.gpjson_last_error(), bytes)


@pytest.mark.skipif(not os.path.exists(_native.LIB_PATH), reason="libgpjson_hip.so not built")
def test_argument_errors_without_device():
    lib = _native.load()
    assert lib.gpjson_ctx_create(0, None) == _native.GPJSON_ERR_INVALID_ARG
    assert lib.gpjson_index_build(None, None, 1, 0, None) == _native.GPJSON_ERR_INVALID_ARG
    assert lib.gpjson_query_file(None, b"x", None, 0, 0, None) == _native.GPJSON_ERR_INVALID_ARG


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(_native, "LIB_PATH", str(tmp_path / "nope.so"))
    monkeypatch.setattr(_native, "_lib", None)
    with pytest.raises(_native.NativeLibraryMissing):
        _native.load()
