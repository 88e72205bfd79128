"""The in-tile member-chain evaluation of the index kernel
(gpjson_amd/csrc/fused_eval.h), built for the CPU with g++: files are cut
into tiles as the kernel cuts them (tiny tiles too, so that lines cross many
tile boundaries), and every line evaluated inside its tile must get the
oracle's spans (the reference's find_value semantics), while every other line
must be on the interpreter's list, each line exactly once."""
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


def byte_bitmap(f: np.ndarray, c: int) -> np.ndarray:
    """Bit per byte of f that equals c, packed LSB-first into u64 words."""
    W = (len(f) + 63) // 64
    bits = np.zeros(max(W, 1) * 64, dtype=np.uint8)
    bits[:len(f)] = f == c
    return np.packbits(bits, bitorder="little").view("<u8").astype(np.uint64)
TILES = (3, 8, 64, 256, 768)


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "libfused.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "gpjson_amd", "csrc"),
                    os.path.join(ROOT, "tests", "harness", "fused_harness.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.fe_run.restype = ctypes.c_int64
    L.fe_run.argtypes = [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                         ctypes.c_void_p, ctypes.c_void_p]
    return L


def chain_keys(path: str):
    from gpjson_amd import compile_path
    ir = compile_path(path).ir
    keys, pc = [], 0
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
        keys.append(bytes(ir[pc:pc + v]))
        pc += v + 1  # the key, then DOWN
    return keys


stats = {"near": 0}


def check(lib, data: bytes, paths, tiles=TILES, max_depth=3):
    from gpjson_amd import compile_path
    f = np.frombuffer(data, dtype=np.uint8)
    fbuf = f if f.size else np.zeros(1, np.uint8)
    cps = [compile_path(p) for p in paths]
    cps = [c for c in cps if c.max_depth <= max_depth]  # the kernel evaluates chains of up to 3 keys
    if not cps:
        return
    L = max(c.max_depth for c in cps)
    nl = O.newline_index(data)
    s = O.string_index(data)
    lb = O.leveled_index(data, s, L)
    W = max((len(f) + 63) // 64, 1)
    nlm, rq, rc = byte_bitmap(f, 0x0A), byte_bitmap(f, 0x22), byte_bitmap(f, 0x3A)
    s1 = s if s.size else np.zeros(1, np.uint64)
    lb1 = lb if lb.size else np.zeros(max(L, 1), np.uint64)
    for cp in cps:
        keys = chain_keys(cp.path)
        assert len(keys) == cp.max_depth
        kb = np.frombuffer(b"".join(keys) or b"\0", dtype=np.uint8)
        kl = np.array([len(k) for k in keys], dtype=np.int32)
        want = O.find_value(data, nl, s, lb, L, cp.ir, 1).reshape(-1, 2)
        for tw in tiles:
            out = np.full((len(nl), 2), -7, dtype=np.int64)
            done = np.zeros(len(nl), dtype=np.uint8)
            cross = np.full(len(nl) + 1, -7, dtype=np.int64)
            near = np.zeros(1, dtype=np.int64)
            n = lib.fe_run(fbuf.ctypes.data, len(f), s1.ctypes.data, lb1.ctypes.data, L, nlm.ctypes.data,
                           rq.ctypes.data, rc.ctypes.data, tw, len(keys), kb.ctypes.data, kl.ctypes.data, len(nl),
                           out.ctypes.data, done.ctypes.data, cross.ctypes.data, near.ctypes.data)
            assert n != -2, (cp.path, tw, "the register key test disagrees with the walk")
            assert n >= 0, (cp.path, tw, "a line was evaluated twice")
            stats["near"] += int(near[0])
            listed = cross[:n]
            # every line exactly once: in a tile, or on the interpreter's list
            assert len(np.unique(listed)) == n and not done[listed].any(), (cp.path, tw)
            assert int(done.sum()) + n == len(nl), (cp.path, tw, int(done.sum()), n, len(nl))
            ev = done.astype(bool)
            if not np.array_equal(out[ev], want[ev]):
                bad = np.flatnonzero(ev & np.any(out != want, axis=1))[:4]
                raise AssertionError(f"{cp.path} tile {tw} words: lines {bad.tolist()} got {out[bad].tolist()} "
                                     f"want {want[bad].tolist()}")


def test_tweets(lib):
    from gpjson_amd import datagen
    before = stats["near"]
    check(lib, datagen.generate("tweets", 91, 2 << 20).tobytes(), TWEET_PATHS)
    assert stats["near"] > before  # the register fast path was exercised


def test_adversarial(lib):
    from gpjson_amd import datagen
    check(lib, datagen.generate("adversarial", 92, 2 << 20).tobytes(), ADV_PATHS, tiles=(8, 256, 768))


@pytest.mark.parametrize("name", ["simple_single_line.ldjson", "not_found_in_current_level.ldjson", "simple.json",
                                  "boundary.json", "escape_carry.txt", "very_long_string.txt"])
def test_fixtures(lib, name):
    check(lib, fixture_bytes(name), ["$.user.lang", "$.user.id", "$.user.foo", "$.apple.lang", "$.a.a", "$.a",
                                     "$.user"])


@pytest.mark.parametrize("seed", range(8))
def test_soup(lib, seed):
    rng = np.random.default_rng(seed)
    alphabet = np.frombuffer(b'ab"\\\n{}[]:, \t', dtype=np.uint8)
    data = bytes(alphabet[rng.integers(0, len(alphabet), 30_000 + seed)])
    check(lib, data, SOUP_PATHS)


@pytest.mark.parametrize("seed", range(6))
def test_objects_soup(lib, seed):
    """Mostly well-formed lines with keys a/b/x/y, nesting, blanks, escapes."""
    rng = np.random.default_rng(100 + seed)
    parts = []
    for _ in range(2000):
        depth = int(rng.integers(0, 4))
        body = '"v"'
        for _ in range(depth):
            k = ["a", "b", "x", "y", 'a\\"', "a b", ""][int(rng.integers(0, 7))]
            sp = " " * int(rng.integers(0, 3))
            body = '{' + sp + f'"{k}"' + sp + ':' + sp + body + sp + (',"z":1' if rng.random() < 0.5 else '') + '}'
        parts.append(body)
    data = ("\n".join(parts)).encode()
    check(lib, data, SOUP_PATHS + ["$.a.b.x", "$.x.y", "$.a\\\"", "$.a\\\".b", "$['a b'].x"])


def test_empty_and_tiny(lib):
    for data in [b"", b"\n", b"{", b'{"a":1}', b'{"a":1}\n', b'{"a" : {"b" : 2 } }\n\n{"a":3}', b' \n \n',
                 b'"a":1', b'{"a":}', b'{"a":{"b":}}\n{"a"::1}']:
        check(lib, data, ["$.a", "$.a.b"])
