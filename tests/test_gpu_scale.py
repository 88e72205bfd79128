"""The plain path at its real size: 4 GiB of seeded tweets, where offsets reach
past 2^31 (the reference's int offsets, result/ResultItem.java:39,45, stop at
2 GiB).  Checked on the whole file:
  - the newline index equals [0] + flatnonzero(data == '\\n') (numpy, no oracle);
  - each headline path's spans, from a plan built for that path alone (one
    index + query pass, as the benchmark runs it), equal the multithreaded
    oracle's on every line, including the lines that start at or above 2^31.
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SIZE = 4 << 30
PATHS = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers"]


def test_four_gib_against_numpy_and_oracle(ctx):
    from gpjson_amd import compile_path, datagen
    from oracle import oracle as O
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    print(f"[scale] generating {SIZE} bytes of tweets", flush=True)
    data = datagen.generate("tweets", 1, SIZE)
    assert data.size == SIZE
    buf = ctx.stage(data)
    cps = [compile_path(p) for p in PATHS]
    L = max(c.max_depth for c in cps)

    print("[scale] newline index against numpy", flush=True)
    want_nl = np.concatenate([np.zeros(1, dtype=np.int64), np.flatnonzero(data == 10).astype(np.int64)])
    high = int(np.searchsorted(want_nl, 1 << 31))  # first line starting at or above 2^31
    assert 0 < high < want_nl.size
    idx = ctx.index(buf, L)
    got_nl = idx.newline_index()
    idx.free()
    assert np.array_equal(got_nl, want_nl)

    print(f"[scale] oracle index on {threads} threads", flush=True)
    nl, s, lb = O.index_mt(data, L, threads)
    assert np.array_equal(nl, want_nl)
    for cp in cps:
        print(f"[scale] {cp.path}: plan pass and oracle", flush=True)
        plan = ctx.plan(buf, [cp.ir])
        plan.run()
        got = plan.spans()[0]
        plan.free()
        want = O.find_value_mt(data, nl, s, lb, L, cp.ir, cp.num_results, threads).reshape(nl.size,
                                                                                            cp.num_results, 2)
        assert got.shape == want.shape
        assert np.array_equal(got[high:], want[high:]), (cp.path, "lines at or above 2^31")
        assert np.array_equal(got, want), cp.path
        # the high lines do carry results with offsets past 2^31
        assert int((got[high:, 0, 0] >= (1 << 31)).sum()) > 0
    buf.free()


def test_more_than_2_31_newlines(ctx):
    """A file of 2^31 + 2^27 newlines (every line empty): the line count and
    newline offsets pass 2^31 inside one group of 65 536 index tiles, where the
    bookkeeping's scan of per-256-tile newline counts must be 64-bit.  Every
    tile holds more newlines than its scratch row, so the rescan path builds
    the whole newline index."""
    n = (1 << 31) + (1 << 27)
    print(f"[scale] {n} newlines", flush=True)
    data = np.full(n, 10, dtype=np.uint8)
    buf = ctx.stage(data)
    del data
    idx = ctx.index(buf, 1)
    assert idx.lines == n + 1
    nl = idx.newline_index()
    idx.free()
    buf.free()
    assert nl.size == n + 1 and nl[0] == 0
    step = 1 << 27
    for a in range(0, n, step):  # nl[k] = k - 1 for k >= 1
        b = min(a + step, n)
        assert np.array_equal(nl[1 + a:1 + b], np.arange(a, b, dtype=np.int64)), a
