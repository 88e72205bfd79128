"""JSONPath compiler against the reference's IR golden vectors.

Expected bytes re-encode JSONPathParserTest.java:35-42 (IRBuilder sequences;
opcode = Opcode ordinal, LEB128 varints, strings as varint length + UTF-8),
plus the benchmark workloads' query shapes (member chains, ranges, deep
chains); unsupported paths from :56-58.
"""
import pytest

from gpjson_amd import _native

pytestmark = pytest.mark.skipif(not __import__("os").path.exists(_native.LIB_PATH),
                                reason="libgpjson_hip.so not built")

END, STORE, UP, DOWN, KEY, INDEX, STREQ = range(7)


def varint(i):
    out = bytearray()
    while i & ~0x7F:
        out.append((i & 0x7F) | 0x80)
        i >>= 7
    out.append(i)
    return bytes(out)


def key(name):
    b = name.encode()
    return bytes([KEY]) + varint(len(b)) + b


def index(i):
    return bytes([INDEX]) + varint(i)


def streq(s):
    b = ('"' + s + '"').encode()
    return bytes([STREQ]) + varint(len(b)) + b


D, U, S, E = bytes([DOWN]), bytes([UP]), bytes([STORE]), bytes([END])

GOLDEN = [
    ("$.user.lang", key("user") + D + key("lang") + D + S + E, 2, 1),
    ("$['user'][\"lang\"]", key("user") + D + key("lang") + D + S + E, 2, 1),
    ("$.user.lang[?(@ == 'nl')]", key("user") + D + key("lang") + D + streq("nl") + S + E, 2, 1),
    ("$[1]", index(1) + D + S + E, 1, 1),
    ("$['us\\'er'][\"lang\"]", key("us\\'er") + D + key("lang") + D + S + E, 2, 1),
    ("$.categoryPath[1:3]", key("categoryPath") + D + index(1) + D + S + U + index(2) + D + S + E, 2, 2),
    ("$.categoryPath[1:3].id", key("categoryPath") + D + index(1) + D + key("id") + D + S + U + U + index(2) + D
     + key("id") + D + S + E, 3, 2),
    ("$.categoryPath[1:4].id", key("categoryPath") + D + index(1) + D + key("id") + D + S + U + U + index(2) + D
     + key("id") + D + S + U + U + index(3) + D + key("id") + D + S + E, 3, 3),
    ("$.user.stats.followers", key("user") + D + key("stats") + D + key("followers") + D + S + E, 3, 1),
    ("$.user.values[999]", key("user") + D + key("values") + D + index(999) + D + S + E, 3, 1),
]

# the same IR written out byte by byte for two of the paths (LEB128 lengths and
# indexes as the reference's IRBuilder encodes them), as an independent check
# of the helpers above
HEX = {
    "$.user.lang": "04 04 75 73 65 72 03 04 04 6c 61 6e 67 03 01 00",
    "$.user.values[999]": "04 04 75 73 65 72 03 04 06 76 61 6c 75 65 73 03 05 e7 07 03 01 00",
}


@pytest.mark.parametrize("path,ir,depth,results", GOLDEN)
def test_golden_ir(path, ir, depth, results):
    from gpjson_amd import compile_path
    cp = compile_path(path)
    assert cp.ir == ir
    assert cp.max_depth == depth
    assert cp.num_results == results


@pytest.mark.parametrize("path", sorted(HEX))
def test_golden_hex(path):
    from gpjson_amd import compile_path
    assert compile_path(path).ir == bytes.fromhex(HEX[path])


@pytest.mark.parametrize("path", ["$.routes[*].legs[*].steps[*].distance.text", "$.meta.view.columns[*].name",
                                  "$.claims.P150[*].mainsnak.property", "$..a", "$[1,2]", "$.a[?(@ > 1)]"])
def test_unsupported(path):
    from gpjson_amd import UnsupportedJSONPathException, compile_path
    with pytest.raises(UnsupportedJSONPathException):
        compile_path(path)


@pytest.mark.parametrize("path", ["$", "$.", "user", "$.a b", "$[]", "$['']", "$[1:]", "$[x]", "$.a[1"])
def test_parse_errors(path):
    from gpjson_amd import JSONPathException, UnsupportedJSONPathException, compile_path
    with pytest.raises(JSONPathException) as ei:
        compile_path(path)
    assert not isinstance(ei.value, UnsupportedJSONPathException)


def test_empty_range_and_deep_chain():
    from gpjson_amd import compile_path
    cp = compile_path("$.a[3:3]")
    assert cp.ir == key("a") + D + S + E and cp.max_depth == 2
    chain = "$.chain" + "".join(f".c{i}" for i in range(1, 20))
    cp = compile_path(chain)
    assert cp.max_depth == 20
