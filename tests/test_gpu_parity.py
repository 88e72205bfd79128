"""GPU parity: every index array and every result span of the HIP path equals the
CPU oracle bit for bit (the oracle is pinned by tests/test_oracle.py).

Inputs: the reference's fixtures, seeded random byte soups aimed at the carry
logic (backslash runs across word / tile boundaries, unterminated strings,
unbalanced brackets, empty and 1-byte inputs), and the tweet / adversarial
generators.  Sizes are chosen so the oracle finishes in seconds.
"""
import json
import os
import sys

import numpy as np
import pytest

from conftest import fixture_bytes
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _p(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def byte_bitmap(data: bytes, c: int) -> np.ndarray:
    """uint64 words with bit i%64 of word i/64 set where data[i] == c (closed form)."""
    f = np.frombuffer(data, dtype=np.uint8)
    W = (len(f) + 63) // 64
    bits = np.zeros(W * 64, dtype=np.uint8)
    bits[:len(f)] = f == c
    return np.packbits(bits, bitorder="little").view("<u8").astype(np.uint64) if W else np.zeros(0, np.uint64)


def level_summary(lb: np.ndarray, levels: int, n: int) -> np.ndarray:
    """Per level, bit k of word s set iff leveled word 64*s+k is not zero (closed form)."""
    W = (n + 63) // 64
    Ws = (W + 63) // 64
    nz = np.zeros((levels, Ws * 64), dtype=np.uint8)
    if W:
        nz[:, :W] = lb.reshape(levels, W) != 0
    return np.packbits(nz, axis=1, bitorder="little").view("<u8").astype(np.uint64).reshape(-1)


def check_index(ctx, data: bytes, levels: int, tag: str = ""):
    _p(f"[parity] index n={len(data)} L={levels} {tag}")
    buf = ctx.stage(data)
    idx = ctx.index(buf, levels, verify=True)
    nl = O.newline_index(data)
    s = O.string_index(data)
    lb = O.leveled_index(data, s, levels)
    assert idx.lines == len(nl)
    assert np.array_equal(idx.newline_index(), nl), "newline index"
    assert np.array_equal(idx.escape_index(), O.escape_index(data)), "escape index"
    assert np.array_equal(idx.quote_index(), O.quote_index(data)), "quote index"
    assert np.array_equal(idx.string_index(), s), "string index"
    assert np.array_equal(idx.colon_bitmap(), byte_bitmap(data, 0x3A)), "colon bitmap"
    assert np.array_equal(idx.raw_quote_bitmap(), byte_bitmap(data, 0x22)), "raw quote bitmap"
    got = idx.leveled_bitmaps()
    if not np.array_equal(got, lb):
        W = (len(data) + 63) // 64
        g2, l2 = got.reshape(levels, W), lb.reshape(levels, W)
        bad = np.argwhere(g2 != l2)[:5]
        raise AssertionError(f"leveled bitmaps differ at (level, word) {bad.tolist()}")
    assert np.array_equal(idx.level_summary(), level_summary(lb, levels, len(data))), "level summaries"
    return buf, idx, nl, s, lb


@pytest.mark.parametrize("name", ["simple.txt", "escape_carry.txt", "very_long_string.txt", "simple.json",
                                  "boundary.json", "even_positions.txt", "odd_positions.txt",
                                  "simple_single_line.ldjson", "not_found_in_current_level.ldjson"])
def test_fixture_indexes(ctx, name):
    check_index(ctx, fixture_bytes(name), 16, name)


def soup(rng, n, alphabet=b'ab"\\\n{}[]:, x'):
    a = np.frombuffer(alphabet, dtype=np.uint8)
    return bytes(a[rng.integers(0, len(a), n)])


@pytest.mark.parametrize("seed", range(8))
def test_random_soup(ctx, seed):
    rng = np.random.default_rng(100 + seed)
    n = int(rng.integers(1, 3_000_000))
    check_index(ctx, soup(rng, n), int(rng.integers(1, 9)), f"soup seed={seed}")


@pytest.mark.parametrize("n", [0, 1, 2, 63, 64, 65, 127, 128, 4095, 4096, 4097, 32767, 32768, 32769, 65536 + 7])
def test_edge_sizes(ctx, n):
    rng = np.random.default_rng(n)
    check_index(ctx, soup(rng, n), 3, f"edge n={n}")


def backslash_runs(rng, n):
    """Long and short backslash runs aligned to word / tile boundaries, quotes after odd and even runs."""
    out = bytearray()
    while len(out) < n:
        k = int(rng.integers(0, 6))
        if k == 0:  # run straddling a 64-byte boundary
            run = int(rng.integers(1, 10))
            lead = (64 - len(out) % 64 - run // 2) % 64
            out += b"a" * lead + b"\\" * run + b'"'
        elif k == 1:  # run straddling a 32 KiB boundary
            run = int(rng.integers(60, 200))
            edge = 32768 - len(out) % 32768
            out += b"x" * max(0, edge - run // 2) + b"\\" * run + (b'"' if rng.integers(0, 2) else b":")
        elif k == 2:
            out += b'{"k":[1,2,"' + b"\\" * int(rng.integers(0, 5)) + b'"],"z":{}}\n'
        else:
            out += soup(rng, int(rng.integers(1, 200)))
    return bytes(out[:n])


def test_short_valid_lines(ctx):
    """Valid NDJSON whose lines are shorter than 64 bytes: every tile holds more
    newlines than it has words (its newline offsets are rescanned from the input
    after the launch) while every speculated carry-in is right."""
    rng = np.random.default_rng(17)
    parts = [b'{"a":1}', b'[]', b'{"k":[1,{"x":"y"}]}', b'"s\\"t"', b'{}', b'3', b'{"b":"\\\\"}']
    out = bytearray()
    while len(out) < 1_200_000:
        out += parts[int(rng.integers(0, len(parts)))] + b"\n"
    check_index(ctx, bytes(out), 3, "short valid lines")


@pytest.mark.parametrize("seed", range(4))
def test_backslash_runs_across_boundaries(ctx, seed):
    rng = np.random.default_rng(200 + seed)
    check_index(ctx, backslash_runs(rng, 1_500_000 + seed * 100_000), 4, f"runs seed={seed}")


@pytest.mark.parametrize("run", [1, 2, 3, 63, 64, 65])
def test_backslash_run_then_quote_at_word_start(ctx, run):
    """A backslash run ending right before a '"' that starts a word, at word,
    64-word row and tile boundaries, with no other backslash near it: the row
    after the run holds none, and its first quote is escaped or not by the
    carry alone (odd runs escape it, even runs do not)."""
    filler = b'{"a":"bc","d":[1,2,{"e":null}],"f":"g h"}\n'
    data = bytearray((filler * (300_000 // len(filler) + 1))[:300_000])
    for b in (64 * 7, 4096 * 3, 4096 * 5 + 64, 32768 * 2, 32768 * 4 + 4096, 49152 * 3, 65536 * 3):
        data[b - run:b] = b"\\" * run
        data[b] = 0x22
    check_index(ctx, bytes(data), 3, f"run={run} then a quote at word starts")


def test_whole_tiles_of_backslashes(ctx):
    # tiles that are entirely backslashes carry the escape state through
    for run in (65536 * 3 + 1, 65536 * 3, 32768 * 5 + 3, 64 * 1000 + 1):
        data = b'{"a":"' + b"\\" * run + b'"}:,[' + b'"q"' * 10 + b"\n"
        check_index(ctx, data, 2, f"all-backslash run={run}")


def test_unterminated_string_and_unbalanced(ctx):
    check_index(ctx, b'{"a": "never closed ' + b"x" * 5000, 2, "open string")
    check_index(ctx, b"]]]]}}}:,:," * 1000 + b"{[{[" * 2000, 8, "unbalanced")
    deep = b"[" * 70 + b"1" + b"]" * 70
    check_index(ctx, deep * 50, 64, "depth 70 with 64 levels")


def spans_equal(ctx, data: bytes, paths, tag=""):
    from gpjson_amd import compile_path
    cps = [compile_path(p) for p in paths]
    L = max(c.max_depth for c in cps)
    R = max(c.num_results for c in cps)
    _p(f"[parity] query n={len(data)} paths={paths} {tag}")
    buf = ctx.stage(data)
    idx = ctx.index(buf, L)
    got = ctx.run_queries(idx, [c.ir for c in cps], R)
    nl, s, lb = O.newline_index(data), O.string_index(data), O.leveled_index(data, O.string_index(data), L)
    for q, c in enumerate(cps):
        ref = O.find_value(data, nl, s, lb, L, c.ir, R).reshape(len(nl), R, 2)
        if not np.array_equal(got[q], ref):
            bad = np.argwhere(got[q] != ref)[:5]
            raise AssertionError(f"{paths[q]}: spans differ at {bad.tolist()}")
    return got


def test_ranges_over_empty_arrays(ctx):
    """`[0:k]` over empty arrays (tweets' "urls" is always [], "hashtags" often):
    the level-end search skips the ']', STORE writes an end of -1, and MOVE_UP out
    of that level ends the line -- in the oracle and on the GPU alike."""
    from gpjson_amd import datagen
    data = (datagen.generate("tweets", 71, 2 << 20).tobytes()
            + b'\n{"a": []}\n{"a": [1]}\n{"a": [ ]}\n{"a": [], "b": [[], [1, 2]]}')
    got = spans_equal(ctx, data, ["$.entities.hashtags[0:2]", "$.entities.urls[0:3]", "$.a[0:2]",
                                  "$.entities.hashtags[0:2].text", "$.b[0:2][0:2]"], "ranges over empty arrays")
    assert (got[1][..., 1] == -1).any()  # the unfound level ends were exercised


KATS = [
    ("$.user.lang", ['"nl"']),
    ("$.user.id", ["8723"]),
    ("$.user.foo", [None]),
    ("$.apple.lang", [None]),
    ("$.user.values[0]", ['"12"']),
    ("$.user.values[2]", ["14"]),
    ("$.user.values[8]", [None]),
    ("$.user.values[0:3]", ['"12"', '"13"', "14"]),
    ("$.user.values[2:4]", ["14", '"15"']),
    ("$.user.lang[?(@ == 'nl')]", ['"nl"']),
]


@pytest.mark.parametrize("path,expected", KATS)
def test_query_function_known_answers(tmp_path, path, expected):
    """QueryFunctionTest.java:86-167 through the full query() member."""
    import gpjson_amd
    f = tmp_path / "simple_single_line.ldjson"
    f.write_bytes(fixture_bytes("simple_single_line.ldjson"))
    _p(f"[kat] {path}")
    with gpjson_amd.GPJSON() as g:
        res = g.query(str(f), path)
        values = [item for line in res[0] for item in line]
    assert values == expected


def test_key_not_found_in_current_level(tmp_path):
    """QueryFunctionTest.java:169-198."""
    import gpjson_amd
    f = tmp_path / "nf.ldjson"
    f.write_bytes(fixture_bytes("not_found_in_current_level.ldjson"))
    with gpjson_amd.GPJSON() as g:
        res = g.query(str(f), "$.a.a")
        count = sum(1 for arr in res for line in arr for v in line if v is not None)
    assert count == 0


def test_multi_query_and_options(tmp_path):
    import gpjson_amd
    f = tmp_path / "s.ldjson"
    f.write_bytes(fixture_bytes("simple_single_line.ldjson"))
    with gpjson_amd.GPJSON() as g:
        res = g.query(str(f), ["$.user.lang", "$.user.values[1:3]", "$.user.foo"])
        assert len(res) == 3 and res.results_per_line == 2
        assert list(res[0][0]) == ['"nl"', None]       # fresh -1 per query
        assert list(res[1][0]) == ['"13"', "14"]
        assert list(res[2][0]) == [None, None]
        with pytest.raises(gpjson_amd.UnsupportedJSONPathException) as e:
            g.query(str(f), "$..lang", {"disable_fallback": True})
        assert not isinstance(e.value, gpjson_amd.FallbackRequired)
        # fallback allowed: the caller is told to evaluate the query itself
        with pytest.raises(gpjson_amd.FallbackRequired) as e:
            g.query(str(f), ["$.user.lang", "$..lang"])
        assert e.value.code == 8 and "(1)" in str(e.value)
        with pytest.raises(gpjson_amd.JSONPathException):
            g.query(str(f), "$.a b")
        with pytest.raises(TypeError):
            g.query(str(f), [])
        with pytest.raises(gpjson_amd.GPJSONException):
            g.query(str(f), ["$.a"] * 256)
        with pytest.raises(gpjson_amd.GPJSONException):
            g.query(str(tmp_path / "missing.ldjson"), "$.a")


TWEET_PATHS = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.entities.hashtags[0]",
               "$.entities.hashtags[1:3].text", "$.lang[?(@ == 'en')]", "$.user.profile.colors.link", "$.extra[2]"]


@pytest.mark.parametrize("size", [1 << 20, 8 << 20])
def test_tweets_all_paths(ctx, size):
    from gpjson_amd import datagen
    data = datagen.generate("tweets", 11, size).tobytes()
    check_index(ctx, data, 7, "tweets")
    spans_equal(ctx, data, TWEET_PATHS, "tweets")


def test_tweets_user_lang_matches_json(ctx):
    from gpjson_amd import compile_path, datagen
    data = datagen.generate("tweets", 12, 2 << 20).tobytes()
    got = spans_equal(ctx, data, ["$.user.lang"])[0]
    lines = data.split(b"\n")
    for i, line in enumerate(lines):
        a, b = got[i, 0]
        if not line.strip():
            assert a == -1
            continue
        obj = json.loads(line)
        if "user" in obj:
            assert data[a:b].decode() == json.dumps(obj["user"]["lang"])


ADV_PATHS = ["$.chain" + "".join(f".c{i}" for i in range(1, 20)), "$.arr[0]", "$.arr[999]", "$.arr[500].v",
             "$.deep.d.d.d", "$.s1", "$.id"]


def test_adversarial_all_paths(ctx):
    from gpjson_amd import datagen
    data = datagen.generate("adversarial", 21, 6 << 20).tobytes()
    check_index(ctx, data, 20, "adversarial L=20")
    check_index(ctx, data, 33, "adversarial L=33")
    got = spans_equal(ctx, data, ADV_PATHS, "adversarial")
    lines = data.split(b"\n")
    checked = 0
    for i, line in enumerate(lines[:50]):
        if not line.strip() or b'"chain"' not in line:
            continue
        obj = json.loads(line)
        a, b = got[2][i, 0]
        assert json.loads(data[a:b]) == obj["arr"][999]
        a, b = got[0][i, 0]
        v = obj["chain"]
        for k in range(1, 20):
            v = v[f"c{k}"]
        assert json.loads(data[a:b]) == v
        checked += 1
    assert checked > 0


def test_rebuild_is_idempotent(ctx):
    from gpjson_amd import datagen
    data = datagen.generate("tweets", 5, 4 << 20).tobytes()
    buf = ctx.stage(data)
    idx = ctx.index(buf, 3)
    a = (idx.newline_index(), idx.string_index(), idx.leveled_bitmaps())
    for _ in range(3):
        idx.rebuild()
        b = (idx.newline_index(), idx.string_index(), idx.leveled_bitmaps())
        assert all(np.array_equal(x, y) for x, y in zip(a, b))


def test_large_against_mt_oracle(ctx):
    """256 MiB of tweets: every array and $.user.lang spans against the multithreaded oracle."""
    from gpjson_amd import compile_path, datagen
    data = datagen.generate("tweets", 31, 256 << 20)
    _p("[parity] 256 MiB tweets vs MT oracle")
    buf = ctx.stage(data)
    cp = compile_path("$.user.lang")
    idx = ctx.index(buf, 2)
    nl, s, lb = O.index_mt(data, 2, 16)
    assert np.array_equal(idx.newline_index(), nl)
    assert np.array_equal(idx.string_index(), s)
    assert np.array_equal(idx.leveled_bitmaps(), lb)
    got = ctx.run_queries(idx, [cp.ir], 1)
    ref = O.find_value_mt(data, nl, s, lb, 2, cp.ir, 1, 16)
    assert np.array_equal(got[0].reshape(-1), ref)


def short_lines(rng, lines: int) -> bytes:
    """NDJSON of `lines` short records; every seventh lacks "a", every eleventh is blank."""
    v = rng.integers(0, 10 ** 6, size=lines)
    out = []
    for i in range(lines):
        if i % 11 == 5:
            out.append("")
        elif i % 7 == 3:
            out.append('{"c":[%d]}' % v[i])
        else:
            out.append('{"a":{"b":%d},"c":[%d,"x%d"]}' % (v[i], i, v[i] % 97))
    return ("\n".join(out) + "\n").encode()


@pytest.mark.parametrize("lines", [(1 << 17) + 13, (1 << 18) + 5, 600001])
def test_many_short_lines(ctx, lines):
    """An interpreter launch of 2^17 work items or more splits them into 2, 4 or 8
    ranges with a counter each: one query (the lane refill) and three queries (the
    claim loop, 3 * lines items) over every line, against the oracle."""
    data = short_lines(np.random.default_rng(lines), lines)
    assert data.count(b"\n") == lines
    spans_equal(ctx, data, ["$.a.b"], f"lines={lines}")
    spans_equal(ctx, data, ["$.a.b", "$.c[1]", "$.c[0]"], f"lines={lines}")


@pytest.mark.parametrize("pairs", [157, 158, 159, 318])
def test_rows_at_the_bracket_list_capacity(ctx, pairs):
    """Every 4096-byte row (a wavefront's 64 words) is one line `{"x":[[],[],...]}`
    of 4 + 2 * pairs brackets, padded with spaces: 318, 320, 322 and 640
    brackets per row, on both sides of the spread list's 320 items."""
    body = '{"x":[' + ",".join(["[]"] * pairs) + "]}"
    line = (body + " " * (4095 - len(body)) + "\n").encode()
    assert len(line) == 4096 and line.count(b"[") + line.count(b"]") + 2 == 4 + 2 * pairs
    data = line * 24
    check_index(ctx, data, 3, f"pairs={pairs}")
    spans_equal(ctx, data, ["$.x[0]", "$.x[100]", f"$.x[{pairs - 1}]"], f"pairs={pairs}")
