"""Host-side mirror of GPJSON's polyglot surface over the C ABI.

Reference interface (paths relative to gpjson/src/main/java/com/koenv/gpjson/):

* ``Polyglot.eval('gpjson', 'jsonpath')`` returns the library object
  (GPJSONLanguage.java:27-34, GPJSONLibrary.java:48-54) -> ``GPJSON()``;
* its member ``query(file, path | [paths], {disable_fallback})``
  (functions/QueryFunction.java:48-265) -> ``GPJSON.query``;
* the returned ``Result[query][line][k]`` -> ``str`` (UTF-8 of the span) or
  ``None`` for the interop null (result/Result.java, ResultArray.java,
  ResultItem.java:32-52, NullValue.java).

Errors map to the reference's exception classes: GPJSONException (device /
runtime failures and bad arguments), JSONPathException (parse errors) and
UnsupportedJSONPathException (paths outside the subset; the reference would
fall back to jayway json-path on the CPU, which this library does not carry).

Lower-level handles (Context, Buffer, Index) expose the index stages for the
parity tests the way the reference's KernelTest exposes its stage classes
(gpjson/src/test/java/com/koenv/gpjson/KernelTest.java:17-45).
"""
from __future__ import annotations

import ctypes
import os
import weakref
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _native as N


class GPJSONException(RuntimeError):
    """Runtime failure (J/GPJSONException.java)."""

    def __init__(self, msg: str, code: int = N.GPJSON_ERR_HIP):
        super().__init__(msg)
        self.code = code


class JSONPathException(GPJSONException):
    """Malformed JSONPath (J/jsonpath/JSONPathException.java)."""


class UnsupportedJSONPathException(JSONPathException):
    """JSONPath outside the supported subset (J/jsonpath/UnsupportedJSONPathException.java)."""


class FallbackRequired(UnsupportedJSONPathException):
    """query() met a path outside the supported subset with fallback allowed: the
    reference answers it with jayway json-path on the CPU (J/functions/
    QueryFunction.java:122-128, 267-314); this library runs nothing and leaves
    that evaluation to the caller's own JSONPath library."""


def _check(rc: int) -> None:
    if rc == N.GPJSON_OK:
        return
    msg = N.last_error()
    if rc == N.GPJSON_FALLBACK_REQUIRED:
        raise FallbackRequired(msg, rc)
    if rc == N.GPJSON_ERR_UNSUPPORTED:
        raise UnsupportedJSONPathException(msg, rc)
    if rc == N.GPJSON_ERR_PARSE:
        raise JSONPathException(msg, rc)
    if rc == N.GPJSON_ERR_INVALID_ARG:
        raise ValueError(msg)
    raise GPJSONException(msg, rc)


class CompiledPath:
    __slots__ = ("path", "ir", "max_depth", "num_results")

    def __init__(self, path: str, ir: bytes, max_depth: int, num_results: int):
        self.path = path
        self.ir = ir
        self.max_depth = max_depth
        self.num_results = num_results

    def __repr__(self) -> str:
        return f"CompiledPath({self.path!r}, ir={self.ir.hex(' ')}, max_depth={self.max_depth})"


def compile_path(path: str) -> CompiledPath:
    """JSONPathParser.compile (J/jsonpath/JSONPathParser.java:20-30)."""
    lib = N.load()
    raw = path.encode("utf-8")
    n = ctypes.c_size_t(0)
    d = ctypes.c_int(0)
    r = ctypes.c_int(0)
    _check(lib.gpjson_compile(raw, None, 0, ctypes.byref(n), ctypes.byref(d), ctypes.byref(r)))
    buf = (ctypes.c_uint8 * max(n.value, 1))()
    _check(lib.gpjson_compile(raw, buf, n.value, ctypes.byref(n), ctypes.byref(d), ctypes.byref(r)))
    return CompiledPath(path, bytes(buf[: n.value]), d.value, r.value)


class Context:
    """One device + one stream (replaces CUDARuntime, J/gpu/CUDARuntime.java:76-94)."""

    def __init__(self, device: int = 0):
        self._lib = N.load()
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_ctx_create(device, ctypes.byref(h)))
        self._h = h
        # handles created on this context; freed before the context itself
        self._children = weakref.WeakSet()

    def _adopt(self, child):
        self._children.add(child)
        return child

    @property
    def handle(self):
        return self._h

    @property
    def num_cus(self) -> int:
        v = ctypes.c_int(0)
        _check(self._lib.gpjson_ctx_num_cus(self._h, ctypes.byref(v)))
        return v.value

    def synchronize(self) -> None:
        _check(self._lib.gpjson_ctx_synchronize(self._h))

    def close(self) -> None:
        if self._h:
            # plans / batches / indexes first, buffers last
            kids = sorted(list(self._children), key=lambda c: isinstance(c, Buffer))
            for c in kids:
                c.free()
            self._lib.gpjson_ctx_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def stage(self, data: Union[bytes, bytearray, memoryview, np.ndarray]) -> "Buffer":
        """Pinned async H2D copy of host bytes."""
        arr = np.frombuffer(memoryview(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else \
            np.ascontiguousarray(data, dtype=np.uint8)
        h = ctypes.c_void_p()
        ptr = arr.ctypes.data_as(ctypes.c_void_p) if arr.size else None
        _check(self._lib.gpjson_buf_from_host(self._h, ptr, arr.size, ctypes.byref(h)))
        return self._adopt(Buffer(self, h))

    def stage_file(self, path: str) -> "Buffer":
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_buf_from_file(self._h, path.encode(), ctypes.byref(h)))
        return self._adopt(Buffer(self, h))

    def wrap_device(self, ptr: int, n: int) -> "Buffer":
        """Wraps n bytes already in this device's memory (e.g. a torch tensor's data_ptr())."""
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_buf_wrap_device(self._h, ctypes.c_void_p(ptr), n, ctypes.byref(h)))
        return self._adopt(Buffer(self, h))

    def index(self, buf: "Buffer", levels: int, verify: bool = False) -> "Index":
        h = ctypes.c_void_p()
        flags = N.GPJSON_BUILD_VERIFY if verify else 0
        _check(self._lib.gpjson_index_build(self._h, buf.handle, levels, flags, ctypes.byref(h)))
        return self._adopt(Index(self, h, buf))

    def index_file(self, path: str, levels: int, verify: bool = False) -> Tuple["Buffer", "Index"]:
        """Reads and indexes a file with the copy of each chunk overlapped with
        indexing the chunks before it; equals stage_file() then index()."""
        hb, hi = ctypes.c_void_p(), ctypes.c_void_p()
        flags = N.GPJSON_BUILD_VERIFY if verify else 0
        _check(self._lib.gpjson_index_file(self._h, path.encode(), levels, flags, ctypes.byref(hb),
                                           ctypes.byref(hi)))
        buf = self._adopt(Buffer(self, hb))
        return buf, self._adopt(Index(self, hi, buf))

    def run_queries(self, idx: "Index", irs: Sequence[bytes], results_per_line: int) -> np.ndarray:
        """int64[nq][lines][R][2] spans for compiled IRs over one index."""
        nq = len(irs)
        bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
        ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
        lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
        lines = idx.lines
        out = np.empty((nq, lines, results_per_line, 2), dtype=np.int64)
        _check(self._lib.gpjson_query_run(self._h, idx.handle, ptrs, lens, nq, results_per_line,
                                          out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def query_gather(self, idx: "Index", irs: Sequence[bytes], results_per_line: int) -> "Gathered":
        """Queries over one index with the result bytes gathered on the device and copied back once."""
        nq = len(irs)
        bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
        ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
        lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_query_gather(self._h, idx.handle, ptrs, lens, nq, results_per_line,
                                             ctypes.byref(h)))
        return Gathered(h, nq)

    def plan(self, buf: "Buffer", irs: Sequence[bytes]) -> "Plan":
        """Index + queries prepared once over a resident buffer (runs allocate nothing)."""
        nq = len(irs)
        bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
        ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
        lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_plan_create(self._h, buf.handle, ptrs, lens, nq, ctypes.byref(h)))
        return self._adopt(Plan(self, h, buf, nq))

    def batch(self, bufs: Sequence["Buffer"], irs_per_file: Sequence[Sequence[bytes]]) -> "Batch":
        """Many resident files: one persistent index launch + one interpreter launch per run."""
        flat = [bytes(ir) for irs in irs_per_file for ir in irs]
        cbufs = [ctypes.create_string_buffer(ir, max(len(ir), 1)) for ir in flat]
        ptrs = (ctypes.c_void_p * len(flat))(*[ctypes.cast(b, ctypes.c_void_p) for b in cbufs])
        lens = (ctypes.c_int64 * len(flat))(*[len(ir) for ir in flat])
        nq = (ctypes.c_int * len(bufs))(*[len(irs) for irs in irs_per_file])
        hb = (ctypes.c_void_p * len(bufs))(*[b.handle for b in bufs])
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_batch_create(self._h, hb, len(bufs), ptrs, lens, nq, ctypes.byref(h)))
        return self._adopt(Batch(self, h, list(bufs), [len(irs) for irs in irs_per_file]))

    def get_config(self) -> dict:
        v = [ctypes.c_int() for _ in range(7)]
        _check(self._lib.gpjson_get_config(self._h, *[ctypes.byref(x) for x in v]))
        keys = ("index_block", "index_wpt", "query_mode", "query_block", "query_group", "query_waves_per_cu",
                "lines_per_group")
        return dict(zip(keys, [x.value for x in v]))

    def set_config(self, index_block: int, index_wpt: int, query_mode: int, query_block: int, query_group: int,
                   query_waves_per_cu: int, lines_per_group: int) -> None:
        """query_mode: 0 line (lane group per line), 3 line with lines of 8 KB or more handed to a
        wavefront each.  Clears the tuning cache."""
        _check(self._lib.gpjson_set_config(self._h, index_block, index_wpt, query_mode, query_block, query_group,
                                           query_waves_per_cu, lines_per_group))

    def set_fusion(self, enable: bool) -> None:
        """In-tile evaluation of single member-chain plans (off by default)."""
        _check(self._lib.gpjson_set_fusion(self._h, 1 if enable else 0))

    def get_fusion(self) -> bool:
        v = ctypes.c_int()
        _check(self._lib.gpjson_get_fusion(self._h, ctypes.byref(v)))
        return bool(v.value)

    def autotune(self, buf: "Buffer", irs: Sequence[bytes], iterations: int = 3) -> str:
        """Times every launch shape on buf and records the fastest in the tuning
        cache for inputs of buf's size bucket at the queries' depth; returns the report."""
        nq = len(irs)
        bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
        ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
        lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
        rep = ctypes.create_string_buffer(16384)
        _check(self._lib.gpjson_autotune(self._h, buf.handle, ptrs, lens, nq, iterations, rep, 16384))
        return rep.value.decode()

    def set_query_order(self, qblock: int) -> None:
        """0: several queries over one file run query after query; B > 0: they take
        turns over blocks of B lines.  Clears the tuning cache."""
        _check(self._lib.gpjson_set_query_order(self._h, qblock))

    def get_query_order(self) -> int:
        v = ctypes.c_int()
        _check(self._lib.gpjson_get_query_order(self._h, ctypes.byref(v)))
        return v.value

    def tune_cache_size(self) -> int:
        return int(self._lib.gpjson_tune_cache_size(self._h))

    def tune_cache_clear(self) -> None:
        _check(self._lib.gpjson_tune_cache_clear(self._h))

    def tune_cache_save(self, path: str) -> None:
        _check(self._lib.gpjson_tune_cache_save(self._h, os.fsencode(path)))

    def tune_cache_load(self, path: str) -> int:
        """Loads the entries of this device's architecture; returns their number."""
        n = ctypes.c_int()
        _check(self._lib.gpjson_tune_cache_load(self._h, os.fsencode(path), ctypes.byref(n)))
        return n.value

    def time_copy(self, nbytes: int, iterations: int = 5) -> float:
        """ms per device-to-device streaming copy of nbytes (reads and writes nbytes)."""
        ms = ctypes.c_double()
        _check(self._lib.gpjson_time_copy(self._h, nbytes, iterations, ctypes.byref(ms)))
        return ms.value

    def time_read(self, nbytes: int, iterations: int = 5) -> float:
        """ms per device streaming read of nbytes (nothing written)."""
        ms = ctypes.c_double()
        _check(self._lib.gpjson_time_read(self._h, nbytes, iterations, ctypes.byref(ms)))
        return ms.value

    def time_plain(self, buf: "Buffer", irs: Sequence[bytes], iterations: int):
        """(ms per index+query pass, index ms, query ms) measured with HIP events on the library stream."""
        nq = len(irs)
        bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
        ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
        lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self._lib.gpjson_time_plain(self._h, buf.handle, ptrs, lens, nq, iterations, ctypes.byref(a),
                                           ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value


class Plan:
    """Prepared plain path: one run = index kernel + one interpreter launch, enqueued without waiting."""

    def __init__(self, ctx: Context, h, buf: "Buffer", nq: int):
        self.ctx, self._h, self.buf, self.nq = ctx, h, buf, nq

    @property
    def lines(self) -> int:
        return int(N.load().gpjson_plan_lines(self._h))

    @property
    def results_per_line(self) -> int:
        return int(N.load().gpjson_plan_results_per_line(self._h))

    def run(self, timed: bool = False) -> None:
        _check(N.load().gpjson_plan_run(self._h, 1 if timed else 0))

    def fused(self):
        """(evaluates lines inside index tiles, lines the last run left to the interpreter or -1)."""
        f, n = ctypes.c_int(), ctypes.c_int64()
        _check(N.load().gpjson_plan_fused(self._h, ctypes.byref(f), ctypes.byref(n)))
        return bool(f.value), n.value

    def timing(self):
        a, b = ctypes.c_double(), ctypes.c_double()
        _check(N.load().gpjson_plan_timing(self._h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def spans(self) -> np.ndarray:
        lib = N.load()
        out = np.empty((self.nq, self.lines, self.results_per_line, 2), dtype=np.int64)
        _check(lib.gpjson_plan_spans(self._h, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def free(self) -> None:
        if self._h and self.ctx._h:
            N.load().gpjson_plan_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Batch:
    """Batched index + query over many resident files (see gpjson_batch_* in include/gpjson.h)."""

    def __init__(self, ctx: Context, h, bufs, nq_per_file):
        self.ctx, self._h, self.bufs, self.nq_per_file = ctx, h, bufs, nq_per_file

    def run(self, graph: bool = True, timed: bool = False) -> None:
        _check(N.load().gpjson_batch_run(self._h, 1 if graph else 0, 1 if timed else 0))

    def autotune(self, iterations: int = 2) -> str:
        """Times the batch's own runs under every launch shape, keeps the fastest
        (and records it in the context's tuning cache); returns the report."""
        rep = ctypes.create_string_buffer(32768)
        _check(N.load().gpjson_batch_autotune(self._h, iterations, rep, 32768))
        return rep.value.decode()

    def get_config(self) -> dict:
        v = [ctypes.c_int() for _ in range(7)]
        _check(N.load().gpjson_batch_get_config(self._h, *[ctypes.byref(x) for x in v]))
        keys = ["index_block", "index_wpt", "query_mode", "query_block", "query_group", "query_waves_per_cu",
                "lines_per_group"]
        return {k: x.value for k, x in zip(keys, v)}

    def query_order(self) -> int:
        return int(N.load().gpjson_batch_query_order(self._h))

    def set_level_split(self, enable: bool) -> None:
        """One index launch per level class (True) or one for every file (False)."""
        _check(N.load().gpjson_batch_set_level_split(self._h, 1 if enable else 0))

    def level_split(self) -> bool:
        return N.load().gpjson_batch_get_level_split(self._h) == 1

    def timing(self):
        a, t = ctypes.c_double(), ctypes.c_double()
        _check(N.load().gpjson_batch_timing(self._h, ctypes.byref(a), ctypes.byref(t)))
        return a.value, t.value

    def lines(self, f: int) -> int:
        v = ctypes.c_int64()
        _check(N.load().gpjson_batch_lines(self._h, f, ctypes.byref(v)))
        return v.value

    def spans(self, f: int, q: int) -> np.ndarray:
        lib = N.load()
        R = lib.gpjson_batch_results_per_line(self._h, f, q)
        if R < 1:
            raise IndexError((f, q))
        out = np.empty((self.lines(f), R, 2), dtype=np.int64)
        _check(lib.gpjson_batch_spans(self._h, f, q, out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out

    def free(self) -> None:
        if self._h and self.ctx._h:
            N.load().gpjson_batch_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Buffer:
    def __init__(self, ctx: Context, h):
        self.ctx = ctx
        self._h = h

    @property
    def handle(self):
        return self._h

    @property
    def size(self) -> int:
        return int(N.load().gpjson_buf_size(self._h))

    def free(self) -> None:
        if self._h and self.ctx._h:
            N.load().gpjson_buf_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Index:
    """Newline / string / leveled (+ escape / quote in verify mode) indexes on the device."""

    def __init__(self, ctx: Context, h, buf: Buffer):
        self.ctx = ctx
        self._h = h
        self.buf = buf  # keeps the input alive

    @property
    def handle(self):
        return self._h

    @property
    def lines(self) -> int:
        return int(N.load().gpjson_index_lines(self._h))

    @property
    def levels(self) -> int:
        return int(N.load().gpjson_index_levels(self._h))

    def rebuild(self) -> None:
        _check(N.load().gpjson_index_rebuild(self._h))

    def export(self, which: int) -> np.ndarray:
        lib = N.load()
        size = ctypes.c_int64(0)
        _check(lib.gpjson_index_export(self._h, which, None, 0, ctypes.byref(size)))
        dtype = np.int64 if which == N.GPJSON_INDEX_NEWLINE else np.uint64
        out = np.empty(size.value // 8, dtype=dtype)
        if size.value:
            _check(lib.gpjson_index_export(self._h, which, out.ctypes.data_as(ctypes.c_void_p), size.value,
                                           ctypes.byref(size)))
        return out

    def newline_index(self) -> np.ndarray:
        return self.export(N.GPJSON_INDEX_NEWLINE)

    def string_index(self) -> np.ndarray:
        return self.export(N.GPJSON_INDEX_STRING)

    def leveled_bitmaps(self) -> np.ndarray:
        return self.export(N.GPJSON_INDEX_LEVELED)

    def escape_index(self) -> np.ndarray:
        return self.export(N.GPJSON_INDEX_ESCAPE)

    def quote_index(self) -> np.ndarray:
        return self.export(N.GPJSON_INDEX_QUOTE)

    def colon_bitmap(self) -> np.ndarray:
        """Every ':' byte (the interpreter's key candidates)."""
        return self.export(N.GPJSON_INDEX_COLON)

    def raw_quote_bitmap(self) -> np.ndarray:
        """Every '"' byte, escaped or not (the interpreter's key-start test)."""
        return self.export(N.GPJSON_INDEX_RAW_QUOTE)

    def level_summary(self) -> np.ndarray:
        """Per level, a bit per leveled-bitmap word that is not zero (the interpreter's skip table)."""
        return self.export(N.GPJSON_INDEX_LEVEL_SUMMARY)

    def free(self) -> None:
        if self._h and self.ctx._h:
            N.load().gpjson_index_free(self._h)
        self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Gathered:
    """Result bytes gathered on the device: per query, spans into the file and
    spans into one contiguous bytes buffer (-1 where there is no result)."""

    def __init__(self, h, nq: int):
        self._lib = N.load()
        self._h = h
        self.number_of_queries = nq
        self.number_of_lines = int(self._lib.gpjson_gathered_lines(h))
        self.results_per_line = int(self._lib.gpjson_gathered_per_line(h))
        p = ctypes.c_void_p()
        n = int(self._lib.gpjson_gathered_bytes(h, ctypes.byref(p)))
        self.bytes = ctypes.string_at(p.value, n) if n else b""

    def _spans(self, fn, q: int) -> np.ndarray:
        p = fn(self._h, q)
        n = self.number_of_lines * self.results_per_line * 2
        if n == 0:
            return np.zeros((0, self.results_per_line, 2), dtype=np.int64)
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.number_of_lines, self.results_per_line, 2).copy()

    def spans(self, q: int) -> np.ndarray:
        return self._spans(self._lib.gpjson_gathered_spans, q)

    def file_spans(self, q: int) -> np.ndarray:
        return self._spans(self._lib.gpjson_gathered_file_spans, q)

    def values(self, q: int) -> List[List[Optional[bytes]]]:
        """[line][k] -> the element's bytes, or None."""
        sp = self.spans(q)
        return [[None if a < 0 else self.bytes[a:b] for a, b in row] for row in sp]

    def timing(self) -> dict:
        a, b, c = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        _check(self._lib.gpjson_gathered_timing(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return {"query_ms": a.value, "gather_ms": b.value, "copy_ms": c.value}

    def free(self) -> None:
        if self._h:
            self._lib.gpjson_gathered_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class ResultItem(Sequence):
    """Results of one query on one line: R entries, each str or None."""

    def __init__(self, res: "Result", q: int, line: int):
        self._res, self._q, self._line = res, q, line

    def __len__(self) -> int:
        return self._res.results_per_line

    def __getitem__(self, k):
        if isinstance(k, slice):
            return [self[i] for i in range(*k.indices(len(self)))]
        if k < 0:
            k += len(self)
        if not 0 <= k < len(self):
            raise IndexError(k)
        return self._res._value(self._q, self._line, k)


class ResultArray(Sequence):
    """One query's results: one ResultItem per line."""

    def __init__(self, res: "Result", q: int):
        self._res, self._q = res, q

    def __len__(self) -> int:
        return self._res.number_of_lines

    def __getitem__(self, line):
        if isinstance(line, slice):
            return [self[i] for i in range(*line.indices(len(self)))]
        if line < 0:
            line += len(self)
        if not 0 <= line < len(self):
            raise IndexError(line)
        return ResultItem(self._res, self._q, line)

    @property
    def spans(self) -> np.ndarray:
        return self._res.spans(self._q)


class Result(Sequence):
    """Result[query][line][k] -> str | None (J/result/Result.java:23-72)."""

    def __init__(self, h):
        self._lib = N.load()
        self._h = h
        self.number_of_queries = int(self._lib.gpjson_result_queries(h))
        self.number_of_lines = int(self._lib.gpjson_result_lines(h))
        self.results_per_line = int(self._lib.gpjson_result_per_line(h))

    def __len__(self) -> int:
        return self.number_of_queries

    def __getitem__(self, q):
        if isinstance(q, slice):
            return [self[i] for i in range(*q.indices(len(self)))]
        if q < 0:
            q += len(self)
        if not 0 <= q < len(self):
            raise IndexError(q)
        return ResultArray(self, q)

    def spans(self, q: int) -> np.ndarray:
        p = self._lib.gpjson_result_spans(self._h, q)
        n = self.number_of_lines * self.results_per_line * 2
        return np.ctypeslib.as_array(p, shape=(n,)).reshape(self.number_of_lines, self.results_per_line, 2).copy()

    def _value(self, q: int, line: int, k: int) -> Optional[str]:
        ptr = ctypes.c_void_p()
        n = ctypes.c_int64(0)
        rc = self._lib.gpjson_result_value(self._h, q, line, k, ctypes.byref(ptr), ctypes.byref(n))
        if rc < 0:
            raise IndexError(N.last_error())
        if rc == 0:
            return None
        return ctypes.string_at(ptr.value, n.value).decode("utf-8", "replace") if n.value else ""

    def close(self) -> None:
        if self._h:
            self._lib.gpjson_result_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class GPJSON:
    """The library object of ``Polyglot.eval('gpjson', 'jsonpath')`` (J/GPJSONLibrary.java:48-98)."""

    MEMBERS = ("query",)

    def __init__(self, device: int = 0):
        self._ctx = Context(device)

    @property
    def context(self) -> Context:
        return self._ctx

    def query(self, file: str, queries: Union[str, Sequence[str]], options: Optional[dict] = None) -> Result:
        """query(file, path | [paths], {disable_fallback}) (J/functions/QueryFunction.java:48-265)."""
        if isinstance(queries, str):
            paths: List[str] = [queries]
        elif isinstance(queries, (list, tuple)):
            if len(queries) > N.GPJSON_MAX_QUERIES:
                raise GPJSONException("Maximum size of queries is 255", N.GPJSON_ERR_INVALID_ARG)
            paths = []
            for q in queries:
                if not isinstance(q, str):
                    raise TypeError("expected queries to be a string or string array")
                paths.append(q)
        else:
            raise TypeError("expected queries to be a string or string array")
        if len(paths) < 1:
            raise TypeError("expected queries to contain at least one query")
        flags = 0
        if options is not None:
            if not isinstance(options, dict):
                raise GPJSONException("Invalid options", N.GPJSON_ERR_INVALID_ARG)
            if options.get("disable_fallback"):
                flags |= N.GPJSON_QUERY_DISABLE_FALLBACK
            # extension: result bytes gathered on the device instead of read from a map of the file
            if options.get("gather"):
                flags |= N.GPJSON_QUERY_GATHER
        arr = (ctypes.c_char_p * len(paths))(*[p.encode("utf-8") for p in paths])
        h = ctypes.c_void_p()
        _check(self._ctx._lib.gpjson_query_file(self._ctx.handle, str(file).encode(), arr, len(paths), flags,
                                                ctypes.byref(h)))
        return Result(h)

    def close(self) -> None:
        self._ctx.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def eval(language: str = "gpjson", source: str = "jsonpath") -> GPJSON:  # noqa: A001 - mirrors Polyglot.eval
    """Polyglot.eval('gpjson', 'jsonpath'); only that source text is accepted (J/GPJSONLanguage.java:29-33)."""
    if language != "gpjson" or source.strip() != "jsonpath":
        raise GPJSONException("Only 'jsonpath' is supported as source", N.GPJSON_ERR_INVALID_ARG)
    return GPJSON()
