"""gpjson_index_file (a file read, copied and indexed in one call): every index
array must equal the build from gpjson_buf_from_file + gpjson_index_build over
the same bytes (which the parity tests pin against the oracle), and query() --
which reads files this way -- must return the spans of that index.

The chunk size is forced small (GPJSON_STAGE_CHUNK_BYTES, read when a context
first stages) so that small files cross many staging-chunk boundaries: chunks
that end inside an index tile, chunks shorter than a tile, and the file's tail.
(The library's pipelined form, which indexes each chunk's tiles as its copy
lands, is switched off; these cases passed with it on.)
"""
import os

import numpy as np
import pytest

from conftest import fixture_bytes

pytestmark = pytest.mark.gpu

# 48 KiB tiles: 1 MiB + 4 KiB + 7 ends chunks mid-tile; 20000 < one tile
CHUNKS = [(1 << 20) + 4096 + 7, 20000, 64 << 20]


def _inputs():
    from gpjson_amd import datagen
    rng = np.random.default_rng(21)
    soup = bytes(np.frombuffer(b'ab"\\\n{}[]:, x', dtype=np.uint8)[rng.integers(0, 13, 300_000)])
    return [
        ("tweets", datagen.generate("tweets", 71, 6 << 20).tobytes(), 3),
        ("adversarial", datagen.generate("adversarial", 72, 6 << 20).tobytes(), 8),
        ("soup", soup, 3),
        ("fixture", fixture_bytes("simple_single_line.ldjson"), 2),
        ("empty", b"", 1),
    ]


def _exports(idx, verify):
    out = {
        "lines": idx.lines,
        "newline": idx.newline_index(),
        "string": idx.string_index(),
        "leveled": idx.leveled_bitmaps(),
        "colon": idx.colon_bitmap(),
        "raw_quote": idx.raw_quote_bitmap(),
        "summary": idx.level_summary(),
    }
    if verify:
        out["escape"] = idx.escape_index()
        out["quote"] = idx.quote_index()
    return out


@pytest.fixture()
def chunk_env():
    old = os.environ.get("GPJSON_STAGE_CHUNK_BYTES")
    yield
    if old is None:
        os.environ.pop("GPJSON_STAGE_CHUNK_BYTES", None)
    else:
        os.environ["GPJSON_STAGE_CHUNK_BYTES"] = old


@pytest.mark.parametrize("chunk", CHUNKS)
def test_index_file_equals_one_launch(tmp_path, chunk_env, chunk):
    from gpjson_amd.api import Context
    os.environ["GPJSON_STAGE_CHUNK_BYTES"] = str(chunk)
    with Context() as ctx:
        for name, data, L in _inputs():
            f = tmp_path / f"{name}.ldjson"
            f.write_bytes(data)
            for verify in (False, True):
                print(f"[staging] chunk={chunk} {name} n={len(data)} L={L} verify={verify}", flush=True)
                buf_a, idx_a = ctx.index_file(str(f), L, verify=verify)
                buf_b = ctx.stage(data)
                idx_b = ctx.index(buf_b, L, verify=verify)
                try:
                    assert buf_a.size == len(data)
                    ea, eb = _exports(idx_a, verify), _exports(idx_b, verify)
                    for k in eb:
                        assert np.array_equal(ea[k], eb[k]), (name, chunk, verify, k)
                finally:
                    idx_a.free()
                    buf_a.free()
                    idx_b.free()
                    buf_b.free()


def test_query_reads_files_through_the_pipeline(tmp_path, chunk_env):
    import gpjson_amd
    from gpjson_amd import compile_path, datagen
    from gpjson_amd.api import Context
    os.environ["GPJSON_STAGE_CHUNK_BYTES"] = str((1 << 20) + 4096 + 7)
    data = datagen.generate("tweets", 73, 5 << 20).tobytes()
    f = tmp_path / "t.ldjson"
    f.write_bytes(data)
    paths = ["$.user.lang", "$.entities.hashtags[0:2].text", "$.place.country_code"]
    print("[staging] query() over a file read in 1 MiB chunks", flush=True)
    with gpjson_amd.GPJSON() as g:
        res = g.query(str(f), paths)
        spans = [res.spans(q) for q in range(len(paths))]
    cps = [compile_path(p) for p in paths]
    with Context() as ctx:
        buf = ctx.stage(data)
        idx = ctx.index(buf, max(c.max_depth for c in cps))
        R = max(c.num_results for c in cps)
        ref = ctx.run_queries(idx, [c.ir for c in cps], R)
        idx.free()
        buf.free()
    for q in range(len(cps)):
        assert np.array_equal(spans[q], ref[q]), paths[q]
