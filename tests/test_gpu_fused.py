"""In-tile evaluation (a prepared plain pass of one member-chain query): the
index kernel evaluates every line inside an index tile from the tile's LDS
masks and the interpreter only the lines that cross tiles.  Its spans must
equal the oracle's and those of the same plan with in-tile evaluation off,
bit for bit, on every tile shape, on inputs whose speculated carry-in breaks
(the second, non-speculating pass must redo the in-tile work), on tiles with
more lines than they evaluate (short lines), and on edge sizes; non-chain
queries and deeper chains fall back to the interpreter alone."""
import numpy as np
import pytest

from conftest import fixture_bytes

pytestmark = pytest.mark.gpu

SHAPES = [(256, 1), (256, 2), (256, 3), (256, 4), (512, 1), (512, 2), (128, 2), (128, 3), (128, 4)]
TWEET_Q = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name",
           "$.nokey", "$.user.nokey"]
ADV_Q = ["$.id", "$.s1", "$.chain.pad", "$.chain.c1.pad", "$.chain.c1.c2", "$.deep.d.d", "$.s2"]


def _oracle(data: bytes, cp):
    from oracle import oracle as O
    nl = O.newline_index(data)
    s = O.string_index(data)
    lb = O.leveled_index(data, s, cp.max_depth)
    return O.find_value(data, nl, s, lb, cp.max_depth, cp.ir, cp.num_results).reshape(nl.size, cp.num_results, 2)


def _plan_spans(ctx, buf, cp, fuse: bool, runs: int = 2):
    prev = ctx.get_fusion()
    ctx.set_fusion(fuse)
    try:
        plan = ctx.plan(buf, [cp.ir])
        for _ in range(runs):
            plan.run()
        out = plan.spans()[0]
        fused, interpreted = plan.fused()
        plan.free()
        # every chain of at most 3 keys is evaluated in tiles when fusion is on
        assert fused == (fuse and _is_short_chain(cp.ir)), (cp.path, fused)
        if fused:
            assert 0 <= interpreted <= out.shape[0], (cp.path, interpreted)
        return out
    finally:
        ctx.set_fusion(prev)


def _is_short_chain(ir: bytes) -> bool:
    pc, d = 0, 0
    while pc < len(ir) and ir[pc] == 4:
        n, sh = 0, 0
        pc += 1
        while True:
            b = ir[pc]
            pc += 1
            n |= (b & 0x7F) << sh
            if not b & 0x80:
                break
            sh += 7
        pc += n
        if pc >= len(ir) or ir[pc] != 3:
            return False
        pc += 1
        d += 1
    return 1 <= d <= 3 and ir[pc:] == b"\x01\x00"


def _check(ctx, name, data: bytes, paths):
    from gpjson_amd import compile_path
    buf = ctx.stage(data)
    try:
        for p in paths:
            cp = compile_path(p)
            print(f"[fused] {name} n={len(data)} {p}", flush=True)
            want = _oracle(data, cp)
            got = _plan_spans(ctx, buf, cp, True)
            assert got.shape == want.shape, (name, p)
            if not np.array_equal(got, want):
                bad = np.flatnonzero(np.any(got != want, axis=(1, 2)))[:4]
                raise AssertionError(f"{name} {p}: lines {bad.tolist()} got {got[bad].tolist()} "
                                     f"want {want[bad].tolist()}")
            assert np.array_equal(_plan_spans(ctx, buf, cp, False), want), (name, p, "fusion off")
    finally:
        buf.free()


def test_tweets_and_adversarial(ctx):
    from gpjson_amd import datagen
    _check(ctx, "tweets", datagen.generate("tweets", 61, 6 << 20).tobytes(), TWEET_Q)
    _check(ctx, "adversarial", datagen.generate("adversarial", 62, 4 << 20).tobytes(), ADV_Q)


@pytest.mark.parametrize("shape", SHAPES)
def test_every_tile_shape(ctx, shape):
    from gpjson_amd import datagen
    cfg = ctx.get_config()
    try:
        ctx.set_config(shape[0], shape[1], cfg["query_mode"], cfg["query_block"], cfg["query_group"],
                       cfg["query_waves_per_cu"], cfg["lines_per_group"])
        _check(ctx, f"tweets {shape}", datagen.generate("tweets", 63, 3 << 20).tobytes(),
               ["$.user.lang", "$.user.stats.followers"])
    finally:
        ctx.set_config(*cfg.values())


def test_broken_speculation_and_short_lines(ctx):
    rng = np.random.default_rng(64)
    soup = bytes(np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)[rng.integers(0, 13, 400_000)])
    _check(ctx, "soup", soup, ["$.a", "$.a.b", "$.x.y.z"])
    short = b"".join(b'{"id":%d,"a":{"b":%d}}\n' % (i % 10, i) for i in range(60_000))
    _check(ctx, "short lines", short, ["$.id", "$.a.b"])
    # lines longer than a tile: no newline in most tiles
    long_lines = b"".join(b'{"pad":"' + b"x" * (100_000 + 7 * i) + b'","id":%d}\n' % i for i in range(12))
    _check(ctx, "long lines", long_lines, ["$.id", "$.pad"])


def test_edges(ctx):
    from gpjson_amd import datagen
    t = datagen.generate("tweets", 65, 200_000).tobytes()
    for n in [0, 1, 2, 63, 64, 65, 4095, 4097, 49151, 49152, 49153, 98304, 100_000]:
        _check(ctx, f"prefix {n}", t[:n], ["$.user.lang", "$.id"])
    for name in ["simple_single_line.ldjson", "not_found_in_current_level.ldjson", "escape_carry.txt"]:
        _check(ctx, name, fixture_bytes(name), ["$.user.lang", "$.user.id", "$.a.a", "$.a"])
    for data in [b"\n", b"{", b'{"a":1}', b'{"a":1}\n', b'{"a" : {"b" : 2 } }\n\n{"a":3}', b'{"a":}',
                 b'{"a":{"b":}}\n{"a"::1}']:
        _check(ctx, repr(data), data, ["$.a", "$.a.b"])


def test_not_fused_queries(ctx):
    """Non-chain queries and chains deeper than three keys take the interpreter
    alone; the plan must still equal the oracle."""
    from gpjson_amd import datagen
    _check(ctx, "tweets", datagen.generate("tweets", 66, 2 << 20).tobytes(),
           ["$.user.profile.colors.background", "$.entities.hashtags[0].text", "$.user.lang[?(@ == 'en')]"])
