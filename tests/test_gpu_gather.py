"""Device-side result gather: the gathered bytes of every element [q][line][k]
equal the input bytes its span covers (the reference's ResultItem decoding,
J/result/ResultItem.java:32-52), null elements stay null, and the query()
member returns the same values with and without gathering.

The spans themselves are pinned against the oracle by test_gpu_parity.py; here
they come from the same interpreter launch, and the gathered copy is checked
against the input with the closed form data[start:end].
"""
import numpy as np
import pytest

from conftest import fixture_bytes

pytestmark = pytest.mark.gpu

PATHS = {
    "tweets": ["$.user.lang", "$.id", "$.place", "$.entities.hashtags[1:3].text", "$.user", "$.lang[?(@ == 'en')]"],
    # $.s2 is the long string (elements of several KB: the whole-wavefront copy)
    "adversarial": ["$.s2", "$.arr[999]", "$.chain.c1.c2", "$.deep"],
    "soup": ["$.a", "$[1]", "$.x.y"],
}


def _inputs():
    from gpjson_amd import datagen
    rng = np.random.default_rng(11)
    soup = bytes(np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)[rng.integers(0, 13, 300_000)])
    return {
        "tweets": datagen.generate("tweets", 61, 2 << 20).tobytes(),
        "adversarial": datagen.generate("adversarial", 62, 2 << 20).tobytes(),
        "soup": soup,
    }


def check_gather(ctx, data: bytes, paths, tag: str):
    from gpjson_amd import compile_path
    print(f"[gather] {tag} n={len(data)}", flush=True)
    cps = [compile_path(p) for p in paths]
    L = max(c.max_depth for c in cps)
    R = max(c.num_results for c in cps)
    buf = ctx.stage(data)
    idx = ctx.index(buf, L)
    plain = ctx.run_queries(idx, [c.ir for c in cps], R)
    g = ctx.query_gather(idx, [c.ir for c in cps], R)
    try:
        expect_off = 0
        for q in range(len(cps)):
            fs, gs = g.file_spans(q), g.spans(q)
            assert np.array_equal(fs, plain[q]), (tag, paths[q], "file spans")
            null = fs[..., 0] < 0
            assert np.array_equal(gs[..., 0] < 0, null), (tag, paths[q], "nulls")
            for line in range(fs.shape[0]):
                for k in range(R):
                    a, b = fs[line, k]
                    if a < 0:
                        continue
                    ga, gb = gs[line, k]
                    # elements are laid out in [q][line][k] order, back to back
                    assert ga == expect_off, (tag, paths[q], line, k, "offset")
                    # an end before the start (a level end the walk never found, on
                    # malformed lines) is an empty element here and in the host
                    # decoder; the reference's ResultItem would fail on it
                    assert g.bytes[ga:gb] == data[a:max(a, b)], (tag, paths[q], line, k, "bytes")
                    expect_off = gb
        assert len(g.bytes) == expect_off
    finally:
        g.free()
        idx.free()
        buf.free()


@pytest.mark.parametrize("name", ["tweets", "adversarial", "soup"])
def test_gather_equals_spans(ctx, name):
    check_gather(ctx, _inputs()[name], PATHS[name], name)


def test_gather_edge_inputs(ctx):
    for tag, data in [("empty", b""), ("one line", b'{"a":1}'), ("no results", b'{"b":1}\n{"b":2}\n'),
                      ("fixture", fixture_bytes("simple_single_line.ldjson"))]:
        check_gather(ctx, data, ["$.a", "$.user.values[0:3]"], tag)


def test_query_member_gather_option(tmp_path):
    import gpjson_amd
    from gpjson_amd import datagen
    f = tmp_path / "t.ldjson"
    f.write_bytes(datagen.generate("tweets", 63, 1 << 20).tobytes())
    paths = ["$.user.lang", "$.entities.hashtags[0:2]", "$.place.country_code"]
    print("[gather] query() with and without gather", flush=True)
    with gpjson_amd.GPJSON() as g:
        a = g.query(str(f), paths)
        b = g.query(str(f), paths, {"gather": True})
        assert len(a) == len(b) and a.number_of_lines == b.number_of_lines
        for q in range(len(a)):
            assert np.array_equal(a.spans(q), b.spans(q))
            for line in range(a.number_of_lines):
                assert list(a[q][line]) == list(b[q][line]), (paths[q], line)
