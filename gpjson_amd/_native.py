"""ctypes binding of libgpjson_hip.so (the C ABI declared in include/gpjson.h).

The library is built in-tree by ``gpjson_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for any product entry point.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
# GPJSON_LIB selects another build of the library (for comparing builds); the
# in-tree build is the default.  There is no fallback: a missing library raises.
LIB_PATH = os.environ.get("GPJSON_LIB") or os.path.join(HERE, "libgpjson_hip.so")

GPJSON_OK = 0
GPJSON_ERR_HIP = 1
GPJSON_ERR_INVALID_ARG = 2
GPJSON_ERR_PARSE = 3
GPJSON_ERR_UNSUPPORTED = 4
GPJSON_ERR_IO = 5
GPJSON_ERR_CAPACITY = 6
GPJSON_ERR_NO_MEMORY = 7
GPJSON_FALLBACK_REQUIRED = 8

GPJSON_INDEX_NEWLINE = 0
GPJSON_INDEX_STRING = 1
GPJSON_INDEX_LEVELED = 2
GPJSON_INDEX_ESCAPE = 3
GPJSON_INDEX_QUOTE = 4
GPJSON_INDEX_COLON = 5
GPJSON_INDEX_RAW_QUOTE = 6
GPJSON_INDEX_LEVEL_SUMMARY = 7

GPJSON_BUILD_VERIFY = 1
GPJSON_QUERY_DISABLE_FALLBACK = 1
GPJSON_QUERY_GATHER = 2
GPJSON_MAX_LEVELS = 64
GPJSON_MAX_QUERIES = 255
GPJSON_NOALLOC_SCRATCH_BYTES = 32 << 20
GPJSON_NOALLOC_MAX_QUERIES = 8
GPJSON_NOALLOC_MAX_IR_BYTES = 1024

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_int64 = ctypes.c_int64
c_size_t = ctypes.c_size_t
c_double = ctypes.c_double
c_char_p = ctypes.c_char_p
P = ctypes.POINTER

# name -> (restype, argtypes); every symbol include/gpjson.h declares
SIGNATURES = {
    "gpjson_last_error": (c_char_p, []),
    "gpjson_ctx_create": (c_int, [c_int, P(c_void_p)]),
    "gpjson_ctx_destroy": (None, [c_void_p]),
    "gpjson_ctx_synchronize": (c_int, [c_void_p]),
    "gpjson_ctx_num_cus": (c_int, [c_void_p, P(c_int)]),
    "gpjson_compile": (c_int, [c_char_p, c_void_p, c_size_t, P(c_size_t), P(c_int), P(c_int)]),
    "gpjson_buf_from_host": (c_int, [c_void_p, c_void_p, c_int64, P(c_void_p)]),
    "gpjson_buf_from_file": (c_int, [c_void_p, c_char_p, P(c_void_p)]),
    "gpjson_buf_wrap_device": (c_int, [c_void_p, c_void_p, c_int64, P(c_void_p)]),
    "gpjson_buf_size": (c_int64, [c_void_p]),
    "gpjson_buf_free": (None, [c_void_p]),
    "gpjson_index_build": (c_int, [c_void_p, c_void_p, c_int, c_int, P(c_void_p)]),
    "gpjson_index_file": (c_int, [c_void_p, c_char_p, c_int, c_int, P(c_void_p), P(c_void_p)]),
    "gpjson_index_rebuild": (c_int, [c_void_p]),
    "gpjson_index_lines": (c_int64, [c_void_p]),
    "gpjson_index_levels": (c_int, [c_void_p]),
    "gpjson_index_export": (c_int, [c_void_p, c_int, c_void_p, c_int64, P(c_int64)]),
    "gpjson_index_free": (None, [c_void_p]),
    "gpjson_query_run": (c_int, [c_void_p, c_void_p, P(c_void_p), P(c_int64), c_int, c_int, c_void_p, c_int64]),
    "gpjson_query_file": (c_int, [c_void_p, c_char_p, P(c_char_p), c_int, c_int, P(c_void_p)]),
    "gpjson_result_queries": (c_int64, [c_void_p]),
    "gpjson_result_lines": (c_int64, [c_void_p]),
    "gpjson_result_per_line": (c_int, [c_void_p]),
    "gpjson_result_spans": (P(c_int64), [c_void_p, c_int64]),
    "gpjson_result_value": (c_int, [c_void_p, c_int64, c_int64, c_int, P(ctypes.c_void_p), P(c_int64)]),
    "gpjson_result_free": (None, [c_void_p]),
    "gpjson_query_gather": (c_int, [c_void_p, c_void_p, P(c_void_p), P(c_int64), c_int, c_int, P(c_void_p)]),
    "gpjson_gathered_lines": (c_int64, [c_void_p]),
    "gpjson_gathered_per_line": (c_int, [c_void_p]),
    "gpjson_gathered_bytes": (c_int64, [c_void_p, P(c_void_p)]),
    "gpjson_gathered_spans": (P(c_int64), [c_void_p, c_int64]),
    "gpjson_gathered_file_spans": (P(c_int64), [c_void_p, c_int64]),
    "gpjson_gathered_timing": (c_int, [c_void_p, P(c_double), P(c_double), P(c_double)]),
    "gpjson_gathered_free": (None, [c_void_p]),
    "gpjson_plan_create": (c_int, [c_void_p, c_void_p, P(c_void_p), P(c_int64), c_int, P(c_void_p)]),
    "gpjson_plan_run": (c_int, [c_void_p, c_int]),
    "gpjson_plan_timing": (c_int, [c_void_p, P(c_double), P(c_double)]),
    "gpjson_plan_fused": (c_int, [c_void_p, P(c_int), P(c_int64)]),
    "gpjson_plan_lines": (c_int64, [c_void_p]),
    "gpjson_plan_results_per_line": (c_int, [c_void_p]),
    "gpjson_plan_spans": (c_int, [c_void_p, c_void_p, c_int64]),
    "gpjson_plan_index": (c_void_p, [c_void_p]),
    "gpjson_plan_free": (None, [c_void_p]),
    "gpjson_batch_create": (c_int, [c_void_p, P(c_void_p), c_int, P(c_void_p), P(c_int64), P(c_int),
                                    P(c_void_p)]),
    "gpjson_batch_run": (c_int, [c_void_p, c_int, c_int]),
    "gpjson_batch_timing": (c_int, [c_void_p, P(c_double), P(c_double)]),
    "gpjson_batch_lines": (c_int, [c_void_p, c_int, P(c_int64)]),
    "gpjson_batch_results_per_line": (c_int, [c_void_p, c_int, c_int]),
    "gpjson_batch_spans": (c_int, [c_void_p, c_int, c_int, c_void_p, c_int64]),
    "gpjson_batch_free": (None, [c_void_p]),
    "gpjson_batch_autotune": (c_int, [c_void_p, c_int, c_char_p, c_size_t]),
    "gpjson_batch_query_order": (c_int, [c_void_p]),
    "gpjson_batch_set_level_split": (c_int, [c_void_p, c_int]),
    "gpjson_batch_get_level_split": (c_int, [c_void_p]),
    "gpjson_set_query_order": (c_int, [c_void_p, c_int]),
    "gpjson_get_query_order": (c_int, [c_void_p, P(c_int)]),
    "gpjson_batch_get_config": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int), P(c_int), P(c_int), P(c_int),
                                        P(c_int)]),
    "gpjson_server_create": (c_int, [c_int, c_int64, c_int, P(c_void_p)]),
    "gpjson_server_add": (c_int, [c_void_p, c_char_p, c_void_p, c_int64]),
    "gpjson_server_admit": (c_int, [c_void_p, c_char_p]),
    "gpjson_server_submit": (c_int, [c_void_p, c_char_p, c_char_p, P(ctypes.c_uint64)]),
    "gpjson_server_wait": (c_int, [c_void_p, ctypes.c_uint64, c_int64, P(c_int64), P(c_int), P(c_double), c_void_p,
                                   c_int64]),
    "gpjson_server_stats": (c_int, [c_void_p, P(c_int64), c_int]),
    "gpjson_server_result": (c_int, [c_void_p, ctypes.c_uint64, c_int64, P(P(c_int64)), P(c_int64), P(c_int),
                                     P(c_double)]),
    "gpjson_server_release": (c_int, [c_void_p, ctypes.c_uint64]),
    "gpjson_server_destroy": (None, [c_void_p]),
    "gpjson_time_copy": (c_int, [c_void_p, c_int64, c_int, P(c_double)]),
    "gpjson_time_read": (c_int, [c_void_p, c_int64, c_int, P(c_double)]),
    "gpjson_get_config": (c_int, [c_void_p, P(c_int), P(c_int), P(c_int), P(c_int), P(c_int), P(c_int), P(c_int)]),
    "gpjson_set_config": (c_int, [c_void_p, c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    "gpjson_set_fusion": (c_int, [c_void_p, c_int]),
    "gpjson_get_fusion": (c_int, [c_void_p, P(c_int)]),
    "gpjson_autotune": (c_int, [c_void_p, c_void_p, P(c_void_p), P(c_int64), c_int, c_int, c_char_p, c_size_t]),
    "gpjson_tune_cache_clear": (c_int, [c_void_p]),
    "gpjson_tune_cache_size": (c_int64, [c_void_p]),
    "gpjson_tune_cache_save": (c_int, [c_void_p, c_char_p]),
    "gpjson_tune_cache_load": (c_int, [c_void_p, c_char_p, P(c_int)]),
    "gpjson_query_noalloc": (c_int, [c_void_p, c_void_p, c_int64, P(c_void_p), P(c_int64), c_int, c_void_p, c_int64,
                                     c_void_p, c_int64, c_void_p]),
    "gpjson_time_plain": (c_int, [c_void_p, c_void_p, P(c_void_p), P(c_int64), c_int, c_int, P(c_double),
                                  P(c_double), P(c_double)]),
}

_lib = None


class NativeLibraryMissing(RuntimeError):
    pass


def load():
    """Loads libgpjson_hip.so; raises NativeLibraryMissing if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise NativeLibraryMissing(
            f"{LIB_PATH} is missing: build it with `python -m gpjson_amd.build` "
            "(there is no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error() -> str:
    return load().gpjson_last_error().decode("utf-8", "replace")
