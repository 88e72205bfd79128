"""CPU oracle pinned against the reference's own fixtures, closed forms and known answers.

Each expectation below is taken from the reference's tests, re-derived
independently of the oracle's code:

* newline index: NewlineIndexTest.java:81-114 (closed form = positions of '\\n'
  with a leading 0, count = #'\\n' + 1), fixtures even/odd_positions.txt;
* escape / quote bitmaps: StringIndexTest.java:109-147 (per-byte carry rule);
* string bitmap: parity of unescaped quotes seen so far (Sequential.java:36-64),
  fixtures simple.txt, escape_carry.txt, very_long_string.txt;
* leveled bitmaps: LeveledBitmapsIndexTest.java (L = 16) on simple.json /
  boundary.json, checked against Python's own json parser for depth;
* query known answers: QueryFunctionTest.java:86-198.
"""
import json
import os

import numpy as np
import pytest

from conftest import fixture_bytes
from oracle import oracle as O


def bits_from_positions(n, positions):
    w = np.zeros((n + 63) // 64, dtype=np.uint64)
    for p in positions:
        w[p // 64] |= np.uint64(1) << np.uint64(p % 64)
    return w


def positions(words, n):
    out = []
    for i in range(n):
        if (int(words[i // 64]) >> (i % 64)) & 1:
            out.append(i)
    return out


@pytest.mark.parametrize("name", ["even_positions.txt", "odd_positions.txt", "simple.txt",
                                  "very_long_string.txt", "boundary.json"])
def test_newline_index_closed_form(name):
    data = fixture_bytes(name)
    nl = O.newline_index(data)
    expect = [0] + [i for i, c in enumerate(data) if c == 0x0A]
    assert nl.tolist() == expect
    assert len(nl) == data.count(b"\n") + 1


def test_newline_fixture_counts():
    # 8192-byte files with a newline at every even / odd position (4096 each)
    ev = O.newline_index(fixture_bytes("even_positions.txt"))
    od = O.newline_index(fixture_bytes("odd_positions.txt"))
    assert len(ev) == 4097 and len(od) == 4097
    assert all(p % 2 == 1 for p in od[1:]) and all(p % 2 == 0 for p in ev[2:])


def _carry_rule(data):
    """StringIndexTest.createExpectedEscapeIndex / createExpectedQuoteIndex."""
    esc, quo = [], []
    carry = 0
    for i, c in enumerate(data):
        if carry == 1:
            esc.append(i)
        if c == 0x22 and carry != 1:
            quo.append(i)
        carry = (carry ^ 1) if c == 0x5C else 0
    return esc, quo


@pytest.mark.parametrize("name", ["simple.txt", "escape_carry.txt", "very_long_string.txt", "boundary.json",
                                  "simple.json", "simple_single_line.ldjson"])
def test_escape_and_quote_index(name):
    data = fixture_bytes(name)
    esc, quo = _carry_rule(data)
    assert np.array_equal(O.escape_index(data), bits_from_positions(len(data), esc))
    assert np.array_equal(O.quote_index(data), bits_from_positions(len(data), quo))


@pytest.mark.parametrize("name", ["simple.txt", "escape_carry.txt", "very_long_string.txt", "boundary.json"])
def test_string_index_is_quote_parity(name):
    data = fixture_bytes(name)
    _, quo = _carry_rule(data)
    qs = set(quo)
    inside, exp = False, []
    for i in range(len(data)):
        if i in qs:
            inside = not inside
        if inside:
            exp.append(i)
    s = O.string_index(data)
    assert np.array_equal(s, bits_from_positions(len(data), exp))


def test_string_index_escape_carry_fixture():
    # escape_carry.txt has an escaped quote (\") outside any string where
    # simple.txt has "..."; it must not toggle the string state, so both files
    # have the same string bitmap.
    a = O.string_index(fixture_bytes("simple.txt"))
    data = fixture_bytes("escape_carry.txt")
    b = O.string_index(data)
    assert np.array_equal(a, b)
    k = data.index(b'\\"')
    assert k + 1 not in positions(b, len(data))
    data = fixture_bytes("very_long_string.txt")
    s = O.string_index(data)
    start = data.index(b'"')
    end = data.rindex(b'"')
    assert positions(s, len(data)) == list(range(start, end))


def _levels_by_json(line: bytes, L: int):
    """Leveled positions derived from a real JSON parse (structure-aware walk)."""
    text = line.decode("utf-8")
    out = {lvl: [] for lvl in range(L)}
    depth = -1
    in_str = False
    esc = False
    # byte offsets: walk bytes, but only ASCII structural chars matter
    for i, c in enumerate(line):
        ch = chr(c)
        if in_str:
            if esc:
                esc = False
            elif ch == "\\":
                esc = True
            elif ch == '"':
                in_str = False
            continue
        if ch == '"':
            in_str = True
        elif ch in "{[":
            depth += 1
            if depth < L:
                out[depth].append(i)
        elif ch in "}]":
            if depth < L:
                out[depth].append(i)
            depth -= 1
        elif ch in ":," and 0 <= depth < L:
            out[depth].append(i)
    json.loads(text)  # the fixture is valid JSON
    return out


@pytest.mark.parametrize("name", ["simple.json", "boundary.json"])
def test_leveled_bitmaps_fixture(name):
    data = fixture_bytes(name)
    L = 16
    s = O.string_index(data)
    lb = O.leveled_index(data, s, L).reshape(L, -1)
    exp = _levels_by_json(data, L)
    for lvl in range(L):
        assert positions(lb[lvl], len(data)) == exp[lvl], lvl


def test_leveled_simple_json_explicit():
    # {"key": "value", "foo":  "bar{]", "array": [123, "56523,3532"]}
    data = fixture_bytes("simple.json")
    s = O.string_index(data)
    lb = O.leveled_index(data, s, 2).reshape(2, -1)
    l0 = positions(lb[0], len(data))
    l1 = positions(lb[1], len(data))
    assert [chr(data[p]) for p in l0] == list("{:,:,:}")
    assert [chr(data[p]) for p in l1] == list("[,]")


def _query(data: bytes, path: str):
    from gpjson_amd.api import compile_path  # product compiler, pinned by test_jsonpath.py

    cp = compile_path(path)
    nl, s, lb, spans = O.query(data, [cp.ir], cp.max_depth, cp.num_results)
    sp = spans[0].reshape(len(nl), cp.num_results, 2)
    vals = []
    for line in sp:
        for a, b in line:
            vals.append(None if a == -1 else data[a:b].decode("utf-8"))
    return vals


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
def test_query_known_answers(path, expected):
    pytest.importorskip("gpjson_amd._native")
    from gpjson_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("compiler library not built")
    assert _query(fixture_bytes("simple_single_line.ldjson"), path) == expected


def test_query_not_found_in_current_level():
    from gpjson_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("compiler library not built")
    vals = _query(fixture_bytes("not_found_in_current_level.ldjson"), "$.a.a")
    assert sum(v is not None for v in vals) == 0


def test_range_over_empty_array_hand_trace():
    """`$.a[0:2]` on `{"a": []}`, traced by hand through K/find_value.cu:21-263:
    KEY a matches the ':' at 4, the cursor moves to 5; MOVE_DOWN finds the end
    of a's value at the '}' (8); MOVE_TO_INDEX 0 skips the blank to the '[' at 6
    and steps to 7; MOVE_DOWN searches level 1 in (8, 8), finds nothing (-1);
    STORE writes (7, -1).  MOVE_UP then makes the cursor -1, which ends the
    line here (the reference would restart at byte 0 of the file)."""
    from gpjson_amd.api import compile_path
    data = b'{"a": []}'
    cp = compile_path("$.a[0:2]")
    nl, s, lb, spans = O.query(data, [cp.ir], cp.max_depth, cp.num_results)
    assert spans[0].reshape(-1).tolist() == [7, -1, -1, -1]
    # the same document with one element: both results from the first branch on
    data = b'{"a": [1]}'
    nl, s, lb, spans = O.query(data, [cp.ir], cp.max_depth, cp.num_results)
    assert spans[0].reshape(-1).tolist() == [7, 8, -1, -1]


def test_oracle_mutation_is_caught():
    """A planted mistake in the rule (treating every quote as unescaped) changes
    the escape_carry fixture's string bitmap, so the checks above can fail."""
    data = fixture_bytes("escape_carry.txt")
    inside, wrong = False, []
    for i, c in enumerate(data):
        if c == 0x22:
            inside = not inside
        if inside:
            wrong.append(i)
    assert not np.array_equal(O.string_index(data), bits_from_positions(len(data), wrong))


def _random_text(rng, n):
    alphabet = np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)
    return bytes(alphabet[rng.integers(0, len(alphabet), n)])


@pytest.mark.parametrize("seed", range(6))
def test_mt_equals_single_thread(seed):
    rng = np.random.default_rng(seed)
    data = _random_text(rng, int(rng.integers(1, 300_000)))
    L = 5
    nl = O.newline_index(data)
    s = O.string_index(data)
    lb = O.leveled_index(data, s, L)
    nl2, s2, lb2 = O.index_mt(data, L, int(rng.integers(1, 9)))
    assert np.array_equal(nl, nl2)
    assert np.array_equal(s, s2)
    assert np.array_equal(lb, lb2)


def test_mt_query_equals_single_thread():
    from gpjson_amd import _native, datagen
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("compiler library not built")
    from gpjson_amd.api import compile_path
    data = datagen.generate("tweets", 7, 1 << 20).tobytes()
    for path in ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers"]:
        cp = compile_path(path)
        nl, s, lb, spans = O.query(data, [cp.ir], cp.max_depth, cp.num_results)
        mt = O.find_value_mt(data, nl, s, lb, cp.max_depth, cp.ir, cp.num_results, 4)
        assert np.array_equal(spans[0], mt)


def test_generated_tweets_answers_match_json():
    """$.user.lang spans on generated lines equal json.dumps of the parsed value."""
    from gpjson_amd import _native, datagen
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("compiler library not built")
    data = datagen.generate("tweets", 3, 256 << 10).tobytes()
    vals = _query(data, "$.user.lang")
    lines = data.split(b"\n")
    # line i > 0 starts at the preceding newline; a final empty line has no result
    for i, line in enumerate(lines):
        if not line.strip():
            assert vals[i] is None
            continue
        obj = json.loads(line)
        if "user" in obj:
            assert vals[i] == json.dumps(obj["user"]["lang"])


_MISSING = object()


def _walk(obj, keys):
    """The value a member chain names in a parsed line, or _MISSING when a key
    is missing or a value on the way is not an object (the interpreter then has
    no structural character at the next level, so it stores nothing)."""
    for k in keys:
        if not isinstance(obj, dict) or k not in obj:
            return _MISSING
        obj = obj[k]
    return obj


@pytest.mark.parametrize("path", ["$.id", "$.place.country_code", "$.user.stats.followers"])
def test_generated_tweets_headline_paths_match_json(path):
    """The other headline paths on generated lines: every span parses (json.loads,
    which allows the span's trailing whitespace) to the value json.loads finds in
    the line; lines whose chain is broken (place is null on most lines) have no
    result.  Both outcomes must occur for country_code."""
    from gpjson_amd import _native, datagen
    if not os.path.exists(_native.LIB_PATH):
        pytest.skip("compiler library not built")
    data = datagen.generate("tweets", 5, 256 << 10).tobytes()
    vals = _query(data, path)
    keys = path[2:].split(".")
    seen = {"value": 0, "none": 0}
    for i, line in enumerate(data.split(b"\n")):
        if not line.strip():
            assert vals[i] is None
            continue
        want = _walk(json.loads(line), keys)
        if want is _MISSING:
            assert vals[i] is None, (i, path, vals[i])
            seen["none"] += 1
        else:
            assert vals[i] is not None and json.loads(vals[i]) == want, (i, path, vals[i], want)
            seen["value"] += 1
    assert seen["value"] > 0
    if path == "$.place.country_code":
        assert seen["none"] > seen["value"] > 0, seen
