"""Work order of several queries over one file's lines: with blocks of B lines
the interpreter maps item r to (query, line) by split_item
(gpjson_amd/csrc/query_common.h), built here for the CPU.  The order must be
block by block, and within a block query by query over the block's lines (the
last block holds the lines left), which also makes it a bijection onto
(query, line) pairs."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "liborder.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "gpjson_amd", "csrc"),
                    os.path.join(ROOT, "tests", "harness", "order_harness.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.split_all.restype = None
    L.split_all.argtypes = [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
    return L


def expected(lines, nq, B):
    out = []
    for start in range(0, lines, B):
        block = range(start, min(start + B, lines))
        for q in range(nq):
            out.extend((q, i) for i in block)
    return out


@pytest.mark.parametrize("lines,nq,B", [(1, 1, 1), (1, 8, 8), (7, 8, 8), (8, 8, 8), (9, 8, 8), (64, 3, 8),
                                        (100, 8, 16), (1000, 5, 64), (1000, 1, 4), (63, 2, 64), (65, 7, 4),
                                        (4097, 8, 8), (0, 8, 8)])
def test_split_item_order(lib, lines, nq, B):
    q = np.zeros(max(lines * nq, 1), dtype=np.int64)
    i = np.zeros_like(q)
    lib.split_all(lines, nq, B, q.ctypes.data, i.ctypes.data)
    got = list(zip(q[:lines * nq].tolist(), i[:lines * nq].tolist()))
    assert got == expected(lines, nq, B)
    assert len(set(got)) == lines * nq
