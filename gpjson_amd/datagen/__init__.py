"""Seeded synthetic NDJSON inputs (see gen.c for the shapes).

``generate(kind, seed, size)`` returns exactly ``size`` bytes as a numpy uint8
array; kind ``"tweets"`` is the social-media-post generator, ``"adversarial"``
the long-string / deep-nesting / long-array generator.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgpjson_datagen.so")
KINDS = {"tweets": 0, "adversarial": 1}
_lib = None


def build() -> str:
    tmp = LIB + ".tmp"
    subprocess.run(["gcc", "-O2", "-fPIC", "-shared", "-std=c11", "-o", tmp, os.path.join(HERE, "gen.c"),
                    "-lpthread", "-lm"], check=True)
    os.replace(tmp, LIB)
    return LIB


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        lib = ctypes.CDLL(LIB)
        lib.gjgen_fill.restype = ctypes.c_int
        lib.gjgen_fill.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _lib = lib
    return _lib


def generate(kind: str, seed: int, size: int, threads: int = 0) -> np.ndarray:
    out = np.empty(size, dtype=np.uint8)
    if size == 0:
        return out
    if threads <= 0:
        threads = min(16, os.cpu_count() or 1)
    rc = _load().gjgen_fill(KINDS[kind], seed, out.ctypes.data_as(ctypes.c_void_p), size, threads)
    if rc != 0:
        raise MemoryError("generator failed")
    return out


def write_file(path: str, kind: str, seed: int, size: int) -> str:
    generate(kind, seed, size).tofile(path)
    return path
