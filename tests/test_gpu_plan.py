"""Prepared plans: several plans over different files and query sets, all
created before any runs, each give the one-shot path's (and the oracle's)
bit-exact spans; a plan keeps the tile shape it was prepared with."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

TWEET_PATHS = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name",
               "$.user.stats.friends", "$.retweet_count", "$.user.profile.colors.background"]
ADV_PATHS = ["$.id", "$.s1", "$.chain.pad", "$.chain.c1.pad", "$.chain.c1.c2", "$.chain.c1.c2.c3.c4", "$.deep.d.d",
             "$.s2"]


@pytest.fixture(scope="module")
def files():
    from gpjson_amd import datagen
    return [
        ("tweets", datagen.generate("tweets", 71, (3 << 20) + 5).tobytes(), TWEET_PATHS[:1]),
        ("adversarial", datagen.generate("adversarial", 72, 2 << 20).tobytes(), ADV_PATHS),
        ("tweets", datagen.generate("tweets", 73, (1 << 20) + 999).tobytes(), TWEET_PATHS),
        ("adversarial", datagen.generate("adversarial", 74, 1 << 20).tobytes(), ADV_PATHS[:3]),
    ]


def _one_shot(ctx, buf, cps, R):
    idx = ctx.index(buf, max(c.max_depth for c in cps))
    out = ctx.run_queries(idx, [c.ir for c in cps], R)
    idx.free()
    return out


def test_plans_created_together(ctx, files):
    from gpjson_amd import compile_path
    bufs, plans, want = [], [], []
    for kind, data, paths in files:
        cps = [compile_path(p) for p in paths]
        R = max(c.num_results for c in cps)
        buf = ctx.stage(data)
        bufs.append(buf)
        want.append(_one_shot(ctx, buf, cps, R))
        plans.append(ctx.plan(buf, [c.ir for c in cps]))
    # interleaved runs, no waits in between
    for _ in range(2):
        for k in (2, 0, 3, 1):
            plans[k].run()
    for k, (kind, data, paths) in enumerate(files):
        print(f"[plan] {kind} {len(data)} bytes, {len(paths)} queries", flush=True)
        got = plans[k].spans()
        assert got.shape == want[k].shape
        assert np.array_equal(got, want[k]), (kind, k)
    # and against the oracle for the first path of each file
    for k, (kind, data, paths) in enumerate(files):
        cp = compile_path(paths[0])
        _, _, _, (want0,) = O.query(data, [cp.ir], cp.max_depth, cp.num_results)
        assert np.array_equal(plans[k].spans()[0].reshape(-1), want0)
    for p in plans:
        p.free()
    for b in bufs:
        b.free()


def test_plan_keeps_tile_shape(ctx, files):
    from gpjson_amd import compile_path
    kind, data, _ = files[2]
    cps = [compile_path(p) for p in TWEET_PATHS[:4]]
    cfg = ctx.get_config()
    buf = ctx.stage(data)
    try:
        ctx.set_config(256, 4, 3, 256, 4, 32, 1)  # long lines to a wavefront, 64 KiB index tiles
        plan = ctx.plan(buf, [c.ir for c in cps])
        ctx.set_config(256, 1, 0, 64, 1, 32, 1)   # other tile size and line mode afterwards
        plan.run()
        got = plan.spans()
        want = _one_shot(ctx, buf, cps, 1)
        assert np.array_equal(got, want)
        plan.run(timed=True)
        index_ms, query_ms = plan.timing()
        assert index_ms > 0 and query_ms > 0
        plan.free()
    finally:
        ctx.set_config(*cfg.values())
        buf.free()
