"""Batched execution equals the oracle and the plain path bit for bit (every
file, every query), both for direct runs and for hipGraph replays with eager
work in between.

Besides tweets and adversarial files of uneven size, the batch holds a byte
soup that is not NDJSON (a speculated carry-in goes wrong, so the batch's
second, non-speculating index pass runs), a file of 9-byte lines (every tile
overflows its newline scratch row, and the file overflows the first newline
array size, so the batch is laid out again), and an empty file."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TWEET_Q = ["$.user.lang", "$.id", "$.place.country_code", "$.user.stats.followers", "$.user.name",
           "$.entities.hashtags[0].text", "$.user.profile.colors.link", "$.lang"]
ADV_Q = ["$.chain.c1.c2.c3", "$.id", "$.arr[0]", "$.arr[999]", "$.deep.d.d", "$.s1",
         "$.chain" + "".join(f".c{i}" for i in range(1, 10)), "$.arr[3].v"]


def _files(seed, nfiles):
    from gpjson_amd import datagen
    rng = np.random.default_rng(seed)
    out = []
    for f in range(nfiles):
        kind = "tweets" if f % 3 else "adversarial"
        size = int(rng.integers(1, 600)) * 4096 + int(rng.integers(0, 4096))
        if f == 5:
            size = 17  # tiny file: one partial tile
        out.append((kind, datagen.generate(kind, 1000 + f, size).tobytes()))
    soup = bytes(np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)[rng.integers(0, 13, 200_000)])
    out.insert(3, ("soup", soup))
    out.insert(7, ("short", b"".join(b'{"id":%d}\n' % (i % 10) for i in range(33_000))))
    out.insert(9, ("empty", b""))
    # lines of 9-40 KB (mode 3 hands them to a wavefront each)
    out.append(("long", b"".join(b'{"id":%d,"pad":"%s","user":{"lang":"l%d","name":"n"},"lang":"x"}\n'
                                 % (k, b"y" * (9000 + 997 * k), k % 7) for k in range(32))))
    return out


@pytest.fixture(scope="module")
def batch_inputs():
    """The files, their compiled queries and the oracle's spans per (file, query)."""
    from gpjson_amd import compile_path
    from oracle import oracle as O
    files = _files(7, 12)
    qsets = [[compile_path(p) for p in (ADV_Q if k == "adversarial" else TWEET_Q)] for k, _ in files]
    ref = []
    for (k, d), qs in zip(files, qsets):
        L = max(c.max_depth for c in qs)
        nl = O.newline_index(d)
        s = O.string_index(d)
        lb = O.leveled_index(d, s, L)
        ref.append([O.find_value(d, nl, s, lb, L, c.ir, c.num_results).reshape(nl.size, c.num_results, 2)
                    for c in qs])
    return files, qsets, ref


@pytest.mark.parametrize("mode,qblock", [(0, 0), (3, 0), (0, 8), (3, 4)])
def test_batch_equals_plain(ctx, mode, qblock, batch_inputs):
    """qblock > 0: the queries of a file take turns over blocks of lines, in the
    batch and in the plain multi-query launches alike."""
    cfg = ctx.get_config()
    try:
        ctx.set_config(cfg["index_block"], cfg["index_wpt"], mode, cfg["query_block"], cfg["query_group"],
                       cfg["query_waves_per_cu"], cfg["lines_per_group"])
        ctx.set_query_order(qblock)
        _batch_equals_plain(ctx, *batch_inputs)
    finally:
        ctx.set_config(*cfg.values())
        ctx.set_query_order(0)


def _batch_equals_plain(ctx, files, qsets, oracle):
    bufs = [ctx.stage(d) for _, d in files]
    print("[batch] building", flush=True)
    batch = ctx.batch(bufs, [[c.ir for c in qs] for qs in qsets])
    # plain references
    plain, plain_lines = [], []
    for (k, d), buf, qs in zip(files, bufs, qsets):
        L = max(c.max_depth for c in qs)
        idx = ctx.index(buf, L)
        plain.append([ctx.run_queries(idx, [c.ir], c.num_results)[0] for c in qs])
        plain_lines.append(idx.lines)
        # all of the file's queries in one plain launch (the work order applies)
        R = max(c.num_results for c in qs)
        together = ctx.run_queries(idx, [c.ir for c in qs], R)
        for q, c in enumerate(qs):
            assert np.array_equal(together[q][:, :c.num_results], oracle[len(plain) - 1][q]), (
                "plain multi-query vs oracle", len(plain) - 1, q, k)
        idx.free()

    for f in range(len(files)):
        for q in range(len(qsets[f])):
            assert np.array_equal(plain[f][q], oracle[f][q]), ("plain vs oracle", f, q, files[f][0])

    def check(tag):
        for f in range(len(files)):
            assert batch.lines(f) == plain_lines[f], (tag, f, files[f][0])
            for q in range(len(qsets[f])):
                got = batch.spans(f, q)
                assert np.array_equal(got, oracle[f][q]), (tag, "vs oracle", f, q, files[f][0])
                assert np.array_equal(got, plain[f][q]), (tag, f, q)

    check("first run")
    batch.run(graph=False)
    check("direct run")
    for r in range(3):
        print(f"[batch] graph replay {r}", flush=True)
        batch.run(graph=True, timed=True)
        ctx.synchronize()
        check(f"graph replay {r}")
        # eager work between replays
        idx = ctx.index(bufs[0], 2)
        idx.free()
    batch.free()


def test_batch_autotune_and_tuning_cache(ctx, batch_inputs, tmp_path):
    """The batch autotuner re-lays the batch out for every candidate shape and
    keeps the fastest: the spans right after it, after a direct run and after
    graph replays equal the oracle's.  Its choice goes into the tuning cache,
    where a second batch of the same files finds it; the cache survives a save,
    clear and load."""
    files, qsets, oracle = batch_inputs
    ctx.tune_cache_clear()
    bufs = [ctx.stage(d) for _, d in files]
    irs = [[c.ir for c in qs] for qs in qsets]
    batch = ctx.batch(bufs, irs)
    try:
        print("[batch] autotune", flush=True)
        report = batch.autotune(1)
        assert "batch best index" in report
        tuned = batch.get_config()
        assert ctx.tune_cache_size() == 1

        def check(tag):
            for f in range(len(files)):
                for q in range(len(qsets[f])):
                    assert np.array_equal(batch.spans(f, q), oracle[f][q]), (tag, f, q, files[f][0], tuned)

        check("after autotune")
        batch.run(graph=False)
        check("direct run")
        for r in range(2):
            batch.run(graph=True)
            ctx.synchronize()
            check(f"graph replay {r}")
        second = ctx.batch(bufs, irs)
        try:
            assert second.get_config() == tuned
        finally:
            second.free()
        path = str(tmp_path / "tune.txt")
        ctx.tune_cache_save(path)
        ctx.tune_cache_clear()
        assert ctx.tune_cache_size() == 0
        assert ctx.tune_cache_load(path) == 1
        assert ctx.tune_cache_size() == 1
        # entries of another architecture are skipped; unsupported shapes,
        # malformed lines and foreign files are refused
        header, entry = open(path).read().splitlines()[:2]
        fields = entry.split()

        def write(name, lines):
            q = str(tmp_path / name)
            with open(q, "w") as fh:
                fh.write("\n".join(lines) + "\n")
            return q

        other = write("other.txt", [header, " ".join(["gfx000"] + fields[1:])])
        assert ctx.tune_cache_load(other) == 0
        bad_shape = write("bad.txt", [header, " ".join(fields[:4] + ["100"] + fields[5:])])
        with pytest.raises(Exception):
            ctx.tune_cache_load(bad_shape)
        with pytest.raises(Exception):
            ctx.tune_cache_load(write("short.txt", [header, " ".join(fields[:6])]))
        with pytest.raises(Exception):
            ctx.tune_cache_load(write("foreign.txt", ["something else", entry]))
        assert ctx.tune_cache_size() == 1
    finally:
        batch.free()
        ctx.tune_cache_clear()
        for b in bufs:
            b.free()


def test_level_split(ctx, batch_inputs):
    """One index launch per level class (the tweets files run the 4-level
    variant, the adversarial files the deepest one) gives the oracle's spans,
    on a direct run and on graph replays, and back to one launch again."""
    files, qsets, oracle = batch_inputs
    bufs = [ctx.stage(d) for _, d in files]
    batch = ctx.batch(bufs, [[c.ir for c in qs] for qs in qsets])
    try:
        for split in (True, False, True):
            batch.set_level_split(split)
            assert batch.level_split() == split
            for tag in ("after the switch", "direct run", "graph replay", "graph replay"):
                if tag == "direct run":
                    batch.run(graph=False)
                elif tag == "graph replay":
                    batch.run(graph=True)
                    ctx.synchronize()
                for f in range(len(files)):
                    for q in range(len(qsets[f])):
                        assert np.array_equal(batch.spans(f, q), oracle[f][q]), (split, tag, f, q, files[f][0])
    finally:
        batch.free()
        for b in bufs:
            b.free()


def test_plain_autotune_cache(ctx):
    """gpjson_autotune records its choice for the input's size bucket and
    depth; plans and one-shot queries that then run with it equal the oracle."""
    from gpjson_amd import compile_path, datagen
    from oracle import oracle as O
    data = datagen.generate("tweets", 71, 3 << 20)
    cp = compile_path("$.user.lang")
    nl = O.newline_index(data.tobytes())
    s = O.string_index(data.tobytes())
    lb = O.leveled_index(data.tobytes(), s, cp.max_depth)
    want = O.find_value(data.tobytes(), nl, s, lb, cp.max_depth, cp.ir, 1).reshape(nl.size, 1, 2)
    ctx.tune_cache_clear()
    buf = ctx.stage(data)
    try:
        report = ctx.autotune(buf, [cp.ir], 1)
        assert "cached for inputs of [2^21, 2^22) bytes at 2 levels" in report, report[-300:]
        assert ctx.tune_cache_size() == 1
        plan = ctx.plan(buf, [cp.ir])
        plan.run()
        assert np.array_equal(plan.spans()[0], want)
        plan.free()
        idx = ctx.index(buf, cp.max_depth)
        assert np.array_equal(ctx.run_queries(idx, [cp.ir], 1)[0], want)
        idx.free()
    finally:
        ctx.tune_cache_clear()
        buf.free()
