"""ctypes front end of the CPU oracle (``libgpjson_oracle.so``).

Test infrastructure only: tests/, ``__graft_entry__.smoke()`` and bench.py's
``cpu_baseline`` leg use it as the checker.  The product package never imports
this module.

Every function mirrors one step of the reference's sequential twin
(``gpjson/src/main/java/com/koenv/gpjson/sequential/Sequential.java``) and
returns numpy arrays laid out exactly like the reference's device buffers.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgpjson_oracle.so")
_lib = None

_u8p = ctypes.POINTER(ctypes.c_uint8)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_i64p = ctypes.POINTER(ctypes.c_int64)


def build() -> None:
    """Compile the oracle with its Makefile (gcc)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def _load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        build()
    lib = ctypes.CDLL(_LIB_PATH)
    lib.gjo_newline_index.restype = ctypes.c_int64
    lib.gjo_newline_index.argtypes = [_u8p, ctypes.c_int64, _i64p, ctypes.c_int64]
    for name in ("gjo_escape_index", "gjo_quote_index", "gjo_string_index"):
        fn = getattr(lib, name)
        fn.restype = None
        fn.argtypes = [_u8p, ctypes.c_int64, _u64p]
    lib.gjo_leveled_index.restype = None
    lib.gjo_leveled_index.argtypes = [_u8p, ctypes.c_int64, _u64p, ctypes.c_int, _u64p]
    lib.gjo_find_value.restype = None
    lib.gjo_find_value.argtypes = [_u8p, ctypes.c_int64, _i64p, ctypes.c_int64, _u64p, _u64p,
                                   ctypes.c_int, _u8p, ctypes.c_int64, ctypes.c_int, _i64p]
    lib.gjo_index_mt.restype = ctypes.c_int64
    lib.gjo_index_mt.argtypes = [_u8p, ctypes.c_int64, ctypes.c_int, ctypes.c_int, _i64p,
                                 ctypes.c_int64, _u64p, _u64p]
    lib.gjo_find_value_mt.restype = None
    lib.gjo_find_value_mt.argtypes = [_u8p, ctypes.c_int64, _i64p, ctypes.c_int64, _u64p, _u64p,
                                      ctypes.c_int, _u8p, ctypes.c_int64, ctypes.c_int, _i64p,
                                      ctypes.c_int]
    _lib = lib
    return lib


def _bytes(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    return np.frombuffer(bytes(data), dtype=np.uint8)


def _p(a: np.ndarray, t):
    return a.ctypes.data_as(t)


def words(n: int) -> int:
    return (n + 63) // 64


def newline_index(data) -> np.ndarray:
    f = _bytes(data)
    lib = _load()
    lines = lib.gjo_newline_index(_p(f, _u8p), f.size, None, 0)
    out = np.empty(lines, dtype=np.int64)
    lib.gjo_newline_index(_p(f, _u8p), f.size, _p(out, _i64p), lines)
    return out


def _bitmap(name: str, data) -> np.ndarray:
    f = _bytes(data)
    out = np.zeros(max(words(f.size), 1), dtype=np.uint64)
    getattr(_load(), name)(_p(f, _u8p), f.size, _p(out, _u64p))
    return out[: words(f.size)]


def escape_index(data) -> np.ndarray:
    return _bitmap("gjo_escape_index", data)


def quote_index(data) -> np.ndarray:
    return _bitmap("gjo_quote_index", data)


def string_index(data) -> np.ndarray:
    return _bitmap("gjo_string_index", data)


def leveled_index(data, string_idx: np.ndarray, levels: int) -> np.ndarray:
    """Level-major ``uint64[levels][W]`` (flattened), as the reference lays it out."""
    f = _bytes(data)
    W = words(f.size)
    out = np.zeros(max(W * levels, 1), dtype=np.uint64)
    s = np.ascontiguousarray(string_idx, dtype=np.uint64)
    if s.size == 0:
        s = np.zeros(1, dtype=np.uint64)
    _load().gjo_leveled_index(_p(f, _u8p), f.size, _p(s, _u64p), levels, _p(out, _u64p))
    return out[: W * levels]


def find_value(data, nl: np.ndarray, string_idx: np.ndarray, leveled: np.ndarray, levels: int,
               ir: bytes, num_results: int) -> np.ndarray:
    """Spans ``int64[lines][num_results][2]`` (flattened) for one query."""
    f = _bytes(data)
    nl = np.ascontiguousarray(nl, dtype=np.int64)
    s = np.ascontiguousarray(string_idx, dtype=np.uint64)
    lb = np.ascontiguousarray(leveled, dtype=np.uint64)
    if s.size == 0:
        s = np.zeros(1, dtype=np.uint64)
    if lb.size == 0:
        lb = np.zeros(1, dtype=np.uint64)
    irb = np.frombuffer(bytes(ir), dtype=np.uint8)
    out = np.empty(nl.size * num_results * 2, dtype=np.int64)
    _load().gjo_find_value(_p(f, _u8p), f.size, _p(nl, _i64p), nl.size, _p(s, _u64p),
                           _p(lb, _u64p), levels, _p(irb, _u8p), irb.size, num_results,
                           _p(out, _i64p))
    return out


def query(data, irs, levels: int, num_results: int):
    """Single-threaded index + every query: (nl, S, LB, [spans per query])."""
    nl = newline_index(data)
    s = string_index(data)
    lb = leveled_index(data, s, levels)
    spans = [find_value(data, nl, s, lb, levels, ir, num_results) for ir in irs]
    return nl, s, lb, spans


def index_mt(data, levels: int, nthreads: int):
    f = _bytes(data)
    W = words(f.size)
    s = np.zeros(max(W, 1), dtype=np.uint64)
    lb = np.zeros(max(W * levels, 1), dtype=np.uint64)
    cap = f.size // 64 + 1024
    lib = _load()
    while True:
        nl = np.empty(cap, dtype=np.int64)
        lines = lib.gjo_index_mt(_p(f, _u8p), f.size, levels, nthreads, _p(nl, _i64p), cap,
                                 _p(s, _u64p), _p(lb, _u64p))
        if lines < 0:
            raise MemoryError("oracle index_mt failed")
        if lines <= cap:
            return nl[:lines], s[:W], lb[: W * levels]
        cap = lines


def find_value_mt(data, nl, string_idx, leveled, levels: int, ir: bytes, num_results: int,
                  nthreads: int) -> np.ndarray:
    f = _bytes(data)
    nl = np.ascontiguousarray(nl, dtype=np.int64)
    s = np.ascontiguousarray(string_idx, dtype=np.uint64)
    lb = np.ascontiguousarray(leveled, dtype=np.uint64)
    if s.size == 0:
        s = np.zeros(1, dtype=np.uint64)
    if lb.size == 0:
        lb = np.zeros(1, dtype=np.uint64)
    irb = np.frombuffer(bytes(ir), dtype=np.uint8)
    out = np.empty(nl.size * num_results * 2, dtype=np.int64)
    _load().gjo_find_value_mt(_p(f, _u8p), f.size, _p(nl, _i64p), nl.size, _p(s, _u64p),
                              _p(lb, _u64p), levels, _p(irb, _u8p), irb.size, num_results,
                              _p(out, _i64p), nthreads)
    return out
