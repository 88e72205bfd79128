"""Every launch shape the autotuner may pick gives the oracle's bit-exact answers:
all index tile shapes, and the interpreter in line mode with every lane-group
size, with and without the long-line launch, on mixed query sets and on member
chains alone and together."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

INDEX_SHAPES = [(256, 1), (256, 2), (256, 3), (256, 4), (512, 1), (512, 2), (128, 2), (128, 3), (128, 4)]
# (mode, block, group); mode 3 hands the adversarial input's lines (4-64 KB) to
# the long-line launch
QUERY_SHAPES = [(0, 256, 1), (0, 128, 1), (0, 256, 4), (0, 64, 16), (0, 256, 64), (3, 256, 1), (3, 256, 4),
                (3, 128, 16)]


@pytest.fixture(scope="module")
def inputs():
    from gpjson_amd import datagen
    rng = np.random.default_rng(5)
    soup = bytes(np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)[rng.integers(0, 13, 700_000)])
    return {
        "tweets": datagen.generate("tweets", 41, 3 << 20).tobytes(),
        "adversarial": datagen.generate("adversarial", 42, 3 << 20).tobytes(),
        "soup": soup,
    }


@pytest.mark.parametrize("block,wpt", INDEX_SHAPES)
def test_index_shapes(ctx, inputs, block, wpt):
    cfg = ctx.get_config()
    try:
        ctx.set_config(block, wpt, 0, 256, 4, 32, 1)
        for name, data in inputs.items():
            print(f"[shapes] index {block}x{wpt} {name}", flush=True)
            buf = ctx.stage(data)
            idx = ctx.index(buf, 5, verify=True)
            s = O.string_index(data)
            assert np.array_equal(idx.newline_index(), O.newline_index(data))
            assert np.array_equal(idx.string_index(), s)
            assert np.array_equal(idx.leveled_bitmaps(), O.leveled_index(data, s, 5))
            idx.free()
            buf.free()
    finally:
        ctx.set_config(*cfg.values())


PATHS = {
    "tweets": ["$.user.lang", "$.place.country_code", "$.entities.hashtags[1]", "$.lang[?(@ == 'en')]",
               "$.user.stats.followers"],
    "adversarial": ["$.arr[999]", "$.chain.c1.c2", "$.arr[2:4]", "$.id"],
    "soup": ["$.a", "$[1]", "$.x.y"],
}


@pytest.mark.parametrize("mode,block,group", QUERY_SHAPES)
def test_query_shapes(ctx, inputs, mode, block, group):
    from gpjson_amd import compile_path
    cfg = ctx.get_config()
    try:
        ctx.set_config(cfg["index_block"], cfg["index_wpt"], mode, block, group, 32, 2)
        for name, data in inputs.items():
            cps = [compile_path(p) for p in PATHS[name]]
            L = max(c.max_depth for c in cps)
            R = max(c.num_results for c in cps)
            print(f"[shapes] query mode={mode} {block}/{group} {name}", flush=True)
            buf = ctx.stage(data)
            idx = ctx.index(buf, L)
            got = ctx.run_queries(idx, [c.ir for c in cps], R)
            nl, s = O.newline_index(data), O.string_index(data)
            lb = O.leveled_index(data, s, L)
            for q, c in enumerate(cps):
                ref = O.find_value(data, nl, s, lb, L, c.ir, R).reshape(len(nl), R, 2)
                assert np.array_equal(got[q], ref), (name, PATHS[name][q])
            idx.free()
            buf.free()
    finally:
        ctx.set_config(*cfg.values())


CHAIN_PATHS = {
    "tweets": ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.profile.colors.background",
               "$.nokey", "$.user.nokey"],
    "adversarial": ["$.id", "$.s1", "$.chain.c1.c2", "$.chain.c1.c2.c3.c4", "$.deep.d.d", "$.s2",
                    "$.chain.c1.c2.c3.c4.c5.c6.c7.c8.c9.c10.c11.c12.c13.c14.c15.c16.c17.c18.c19"],
    "soup": ["$.a", "$.a.b", "$.x.y.z"],
}


@pytest.mark.parametrize("block,group,wpc", [(64, 1, 8), (256, 1, 16), (256, 4, 16), (256, 16, 8), (128, 64, 16)])
def test_member_chains(ctx, inputs, block, group, wpc):
    """Member-chain queries equal the oracle, one query per launch and all
    queries in one launch."""
    from gpjson_amd import compile_path
    cfg = ctx.get_config()
    try:
        ctx.set_config(cfg["index_block"], cfg["index_wpt"], 0, block, group, wpc, 1)
        for name, data in inputs.items():
            cps = [compile_path(p) for p in CHAIN_PATHS[name]]
            L = max(c.max_depth for c in cps)
            print(f"[shapes] chain {block}/{group}/{wpc} {name}", flush=True)
            buf = ctx.stage(data)
            idx = ctx.index(buf, L)
            together = ctx.run_queries(idx, [c.ir for c in cps], 1)
            nl, s = O.newline_index(data), O.string_index(data)
            lb = O.leveled_index(data, s, L)
            for q, c in enumerate(cps):
                ref = O.find_value(data, nl, s, lb, L, c.ir, 1).reshape(len(nl), 1, 2)
                alone = ctx.run_queries(idx, [c.ir], 1)[0]
                assert np.array_equal(alone, ref), (name, c.path)
                assert np.array_equal(together[q], ref), (name, c.path)
            idx.free()
            buf.free()
    finally:
        ctx.set_config(*cfg.values())
