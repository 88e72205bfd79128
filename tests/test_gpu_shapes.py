"""Every launch shape the autotuner may pick gives the oracle's bit-exact answers:
all index tile shapes, and the interpreter in tiled and line mode with every
lane-group size."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

INDEX_SHAPES = [(256, 1), (256, 2), (256, 4), (512, 1), (512, 2), (128, 2), (128, 4)]
QUERY_SHAPES = [(1, 256, 1), (1, 256, 4), (1, 128, 16), (1, 256, 64), (0, 256, 1), (0, 256, 4), (0, 64, 16),
                (0, 256, 64)]


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
        ctx.set_config(block, wpt, 1, 256, 4, 32, 1)
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


@pytest.mark.parametrize("tiled,block,group", QUERY_SHAPES)
def test_query_shapes(ctx, inputs, tiled, block, group):
    from gpjson_amd import compile_path
    cfg = ctx.get_config()
    try:
        ctx.set_config(cfg["index_block"], cfg["index_wpt"], tiled, block, group, 32, 2)
        for name, data in inputs.items():
            cps = [compile_path(p) for p in PATHS[name]]
            L = max(c.max_depth for c in cps)
            R = max(c.num_results for c in cps)
            print(f"[shapes] query tiled={tiled} {block}/{group} {name}", flush=True)
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
