"""Word-level bit arithmetic of the index kernel (gpjson_amd/csrc/bitops.h), built
for the CPU with g++ and checked against byte-serial definitions.

* escape_word: all 2^16 low-bit patterns x carry-in, plus structured random
  64-bit words, against the reference's per-byte carry rule
  (StringIndexTest.java:109-127);
* escape_carry_out, prefix_xor: against byte loops;
* classify64: against numpy byte compares on random bytes (all 256 values);
* leveled_word: against the per-byte depth rule (Sequential.java:84-100).
"""
import ctypes
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    out = tmp_path_factory.mktemp("harness") / "libbitops.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", os.path.join(ROOT, "gpjson_amd", "csrc"),
                    os.path.join(ROOT, "tests", "harness", "bitops_harness.cpp"), "-o", str(out)], check=True)
    L = ctypes.CDLL(str(out))
    L.bh_escape_word.restype = ctypes.c_uint64
    L.bh_escape_word.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    L.bh_escape_carry_out.restype = ctypes.c_uint32
    L.bh_escape_carry_out.argtypes = [ctypes.c_uint64, ctypes.c_uint32]
    L.bh_prefix_xor.restype = ctypes.c_uint64
    L.bh_prefix_xor.argtypes = [ctypes.c_uint64]
    L.bh_classify64.restype = None
    L.bh_classify64.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.bh_leveled_word.restype = ctypes.c_int64
    L.bh_leveled_word.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    L.bh_leveled_by_brackets.restype = None
    L.bh_leveled_by_brackets.argtypes = [ctypes.c_uint64] * 3 + [ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                                                                   ctypes.c_void_p]
    return L


def escape_ref(B: int, p: int):
    E, c = 0, p
    for i in range(64):
        if c:
            E |= 1 << i
        c = (c ^ 1) if (B >> i) & 1 else 0
    return E, c


def escape_ref_vec(Bs: np.ndarray, p: int):
    """Vectorized per-bit rule over many words (uint64 array)."""
    E = np.zeros_like(Bs)
    c = np.full(Bs.shape, p, dtype=np.uint64)
    one = np.uint64(1)
    for i in range(64):
        E |= c << np.uint64(i)
        b = (Bs >> np.uint64(i)) & one
        c = np.where(b == one, c ^ one, np.uint64(0))
    return E, c


def test_escape_exhaustive_16bit(lib):
    Bs = np.arange(1 << 16, dtype=np.uint64)
    for p in (0, 1):
        E, c = escape_ref_vec(Bs, p)
        got = np.fromiter((lib.bh_escape_word(int(b), p) for b in Bs), dtype=np.uint64, count=Bs.size)
        assert np.array_equal(got, E)
        # carries out of 16-bit patterns (upper 48 bits zero) are 0 except via the closed form
        cout = np.fromiter((lib.bh_escape_carry_out(int(b), p) for b in Bs[:4096]), dtype=np.uint64, count=4096)
        assert np.array_equal(cout, c[:4096])


def test_escape_structured_random(lib):
    rng = np.random.default_rng(1)
    words = []
    for _ in range(20000):
        w = 0
        pos = 0
        while pos < 64:
            gap = int(rng.integers(0, 6))
            run = int(rng.integers(0, 12))
            pos += gap
            for k in range(run):
                if pos + k < 64:
                    w |= 1 << (pos + k)
            pos += run
        words.append(w)
    words += [(1 << 64) - 1, (1 << 64) - 2, (1 << 63), 1, 0x8000000000000001, 0xFFFF0000FFFF0000]
    for w in words:
        for p in (0, 1):
            E, c = escape_ref(w, p)
            assert lib.bh_escape_word(w, p) == E, (hex(w), p)
            assert lib.bh_escape_carry_out(w, p) == c, (hex(w), p)


def test_prefix_xor(lib):
    rng = np.random.default_rng(2)
    for _ in range(5000):
        q = int(rng.integers(0, 1 << 63)) | (int(rng.integers(0, 2)) << 63)
        acc, ref = 0, 0
        for i in range(64):
            acc ^= (q >> i) & 1
            ref |= acc << i
        assert lib.bh_prefix_xor(q) == ref


def test_classify64(lib):
    rng = np.random.default_rng(3)
    special = np.frombuffer(b'\\"\n{}[]:,', dtype=np.uint8)
    out = np.zeros(7, dtype=np.uint64)
    for t in range(4000):
        if t % 2:
            b = rng.integers(0, 256, 64).astype(np.uint8)
        else:
            b = np.where(rng.random(64) < 0.5, special[rng.integers(0, special.size, 64)],
                         rng.integers(0, 256, 64)).astype(np.uint8)
        lib.bh_classify64(b.ctypes.data, out.ctypes.data)
        def m(mask):
            return int(np.packbits(mask[::-1].astype(np.uint8)).view(">u8")[0]) if False else \
                sum(1 << i for i in range(64) if mask[i])
        assert int(out[0]) == m(b == 0x5C)
        assert int(out[1]) == m(b == 0x22)
        assert int(out[2]) == m(b == 0x0A)
        assert int(out[3]) == m((b == 0x7B) | (b == 0x5B))
        assert int(out[4]) == m((b == 0x7D) | (b == 0x5D))
        assert int(out[5]) == m((b == 0x3A) | (b == 0x2C))
        assert int(out[6]) == m(b == 0x3A)


def leveled_ref(O, C, P, d, lo, L):
    out = [0] * 8
    for i in range(64):
        bit = 1 << i
        if O & bit:
            d += 1
            if lo <= d < lo + 8 and d < L:
                out[d - lo] |= bit
        elif C & bit:
            if lo <= d < lo + 8 and d < L:
                out[d - lo] |= bit
            d -= 1
        elif P & bit:
            if lo <= d < lo + 8 and d < L and d >= 0:
                out[d - lo] |= bit
    return out, d


def test_leveled_word(lib):
    rng = np.random.default_rng(4)
    out = np.zeros(8, dtype=np.uint64)
    for _ in range(20000):
        kinds = rng.integers(0, 5, 64)  # 0,1: none, 2: open, 3: close, 4: sep
        O = sum(1 << i for i in range(64) if kinds[i] == 2)
        C = sum(1 << i for i in range(64) if kinds[i] == 3)
        P = sum(1 << i for i in range(64) if kinds[i] == 4)
        d = int(rng.integers(-3, 12))
        lo = int(rng.choice([0, 8]))
        L = int(rng.integers(1, 17))
        dd = lib.bh_leveled_word(O, C, P, d, lo, L, out.ctypes.data)
        ref, rd = leveled_ref(O, C, P, d, lo, L)
        assert dd == rd
        assert [int(x) for x in out] == ref


def test_leveled_by_brackets(lib):
    """Bracket-at-a-time assembly (each bracket's level from popcounts, the
    separators after it up to the next bracket) equals the per-byte rule."""
    rng = np.random.default_rng(5)
    out = np.zeros(8, dtype=np.uint64)
    for it in range(20000):
        # dense and sparse bracket mixes, with runs that nest several levels
        probs = [0.5, 0.1, 0.15, 0.15, 0.1] if it % 2 else [0.8, 0.05, 0.05, 0.05, 0.05]
        kinds = rng.choice(5, 64, p=probs)  # 0,1: none, 2: open, 3: close, 4: sep
        O = sum(1 << i for i in range(64) if kinds[i] == 2)
        C = sum(1 << i for i in range(64) if kinds[i] == 3)
        P = sum(1 << i for i in range(64) if kinds[i] == 4)
        d = int(rng.integers(-3, 12))
        lo = int(rng.choice([0, 2, 8]))
        L = int(rng.integers(1, 17))
        lib.bh_leveled_by_brackets(O, C, P, d, lo, L, out.ctypes.data)
        ref, _ = leveled_ref(O, C, P, d, lo, L)
        assert [int(x) for x in out] == ref, (it, hex(O), hex(C), hex(P), d, lo, L)
