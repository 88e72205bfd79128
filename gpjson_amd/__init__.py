"""gpjson_amd -- MI355X (gfx950) implementation of GPJSON's NDJSON index + query path.

``gpjson_amd.GPJSON().query(file, "$.user.lang")`` mirrors
``Polyglot.eval('gpjson', 'jsonpath').query(file, path)``; the native library
``libgpjson_hip.so`` (C ABI in include/gpjson.h) does all the work.
"""
from .api import (  # noqa: F401
    GPJSON,
    Buffer,
    CompiledPath,
    Context,
    GPJSONException,
    Index,
    JSONPathException,
    Result,
    ResultArray,
    ResultItem,
    UnsupportedJSONPathException,
    FallbackRequired,
    compile_path,
    eval,
)
from ._native import NativeLibraryMissing  # noqa: F401

__all__ = [
    "GPJSON", "Buffer", "CompiledPath", "Context", "GPJSONException", "Index", "JSONPathException",
    "Result", "ResultArray", "ResultItem", "UnsupportedJSONPathException", "FallbackRequired", "compile_path", "eval",
    "NativeLibraryMissing",
]
