"""The member-chain evaluator's state machine (gpjson_amd/csrc/chain_step.h),
built for the CPU with g++ and run with groups of 1, 4 and 16 lanes, against
the oracle's interpreter (the reference's find_value semantics) on tweets,
adversarial lines, fixtures and random structural soups."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import fixture_bytes
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

TWEET_PATHS = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name",
               "$.user.stats.friends", "$.retweet_count", "$.user.profile.colors.background", "$.user.id",
               "$.place.bounding_box.type", "$.lang", "$.nokey", "$.user.nokey", "$.entities"]
ADV_PATHS = ["$.id", "$.s1", "$.chain.pad", "$.chain.c1.pad", "$.chain.c1.c2", "$.chain.c1.c2.c3.c4", "$.deep.d.d",
             "$.s2", "$.chain.c1.c2.c3.c4.c5.c6.c7.c8.c9.c10.c11.c12.c13.c14.c15.c16.c17.c18.c19", "$.arr"]
SOUP_PATHS = ["$.a", "$.a.b", "$.x.y.z", "$.b", "$['a b']"]


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "libchain.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "gpjson_amd", "csrc"),
                    os.path.join(ROOT, "tests", "harness", "chain_harness.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.ch_run.restype = ctypes.c_int64
    L.ch_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int64]
    return L


def byte_bitmap(f: np.ndarray, c: int) -> np.ndarray:
    W = (len(f) + 63) // 64
    bits = np.zeros(max(W, 1) * 64, dtype=np.uint8)
    bits[:len(f)] = f == c
    return np.packbits(bits, bitorder="little").view("<u8").astype(np.uint64)


def is_chain(ir: bytes) -> bool:
    pc, d = 0, 0
    while pc < len(ir) and ir[pc] == 4:
        pc += 1
        v, sh = 0, 0
        while True:
            b = ir[pc]
            pc += 1
            v |= (b & 0x7F) << sh
            if not b & 0x80:
                break
            sh += 7
        pc += v
        if pc >= len(ir) or ir[pc] != 3:
            return False
        pc += 1
        d += 1
    return d > 0 and ir[pc:] == b"\x01\x00"


GROUPS = (1, 4, 16)


def check(lib, data: bytes, paths, levels=None):
    from gpjson_amd import compile_path
    f = np.frombuffer(data, dtype=np.uint8)
    cps = [compile_path(p) for p in paths]
    L = levels or max(c.max_depth for c in cps)
    nl = O.newline_index(data)
    s = O.string_index(data)
    lb = O.leveled_index(data, s, L)
    W = (len(f) + 63) // 64
    rc, rq = byte_bitmap(f, 0x3A), byte_bitmap(f, 0x22)
    s1 = s if s.size else np.zeros(1, np.uint64)
    lb1 = lb if lb.size else np.zeros(max(L, 1), np.uint64)
    fbuf = f if f.size else np.zeros(1, np.uint8)
    for cp in cps:
        assert is_chain(cp.ir), cp.path
        ir = np.frombuffer(cp.ir, dtype=np.uint8)
        want = O.find_value(data, nl, s, lb, L, cp.ir, 1)
        for G in GROUPS:
            got = np.full(len(nl) * 2, -7, dtype=np.int64)
            steps = lib.ch_run(fbuf.ctypes.data, len(f), nl.ctypes.data, len(nl), s1.ctypes.data, lb1.ctypes.data,
                               max(W, 1), L, rc.ctypes.data, rq.ctypes.data, ir.ctypes.data, 1, got.ctypes.data,
                               G, 1 << 22)
            assert steps >= 0, cp.path
            if not np.array_equal(got, want):
                bad = np.flatnonzero(got != want)[:4]
                raise AssertionError(f"{cp.path} G={G}: differs at {bad.tolist()}: got {got[bad]} want {want[bad]}")


def test_tweets(lib):
    from gpjson_amd import datagen
    check(lib, datagen.generate("tweets", 77, 3 << 20).tobytes(), TWEET_PATHS)


def test_adversarial(lib):
    from gpjson_amd import datagen
    check(lib, datagen.generate("adversarial", 78, 3 << 20).tobytes(), ADV_PATHS)


@pytest.mark.parametrize("name", ["simple_single_line.ldjson", "not_found_in_current_level.ldjson", "simple.json",
                                  "boundary.json", "escape_carry.txt", "very_long_string.txt"])
def test_fixtures(lib, name):
    check(lib, fixture_bytes(name), ["$.user.lang", "$.user.id", "$.user.foo", "$.apple.lang", "$.a.a", "$.a",
                                     "$.user"], levels=4)


@pytest.mark.parametrize("seed", range(12))
def test_soup(lib, seed):
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b'ab"\\\n{}[]:, \t', dtype=np.uint8)
    data = bytes(alphabet[rng.integers(0, len(alphabet), 40_000 + seed)])
    check(lib, data, SOUP_PATHS, levels=4)


@pytest.mark.parametrize("seed", range(6))
def test_objects_soup(lib, seed):
    """Mostly well-formed lines with keys a/b/x/y, nesting, blanks, escapes."""
    rng = np.random.default_rng(100 + seed)
    parts = []
    for _ in range(3000):
        depth = int(rng.integers(0, 4))
        body = '"v"'
        for _ in range(depth):
            k = ["a", "b", "x", "y", 'a\\"', "a b", ""][int(rng.integers(0, 7))]
            sp = " " * int(rng.integers(0, 3))
            body = '{' + sp + f'"{k}"' + sp + ':' + sp + body + sp + (',"z":1' if rng.random() < 0.5 else '') + '}'
        parts.append(body)
    data = ("\n".join(parts)).encode()
    check(lib, data, SOUP_PATHS + ["$.a.b.x", "$.x.y", "$.a\\\"", "$.a\\\".b"], levels=4)


def test_empty_and_tiny(lib):
    for data in [b"", b"\n", b"{", b'{"a":1}', b'{"a":1}\n', b'{"a" : {"b" : 2 } }\n\n{"a":3}', b' \n \n']:
        check(lib, data, ["$.a", "$.a.b"], levels=2)
