"""Query serving over resident indexes (gpjson_server_* in include/gpjson.h).

Server keeps registered files and their indexes resident in HBM within a byte
budget (least-recently-used files are evicted; the host copy stays) and
coalesces every pending query on resident files into one interpreter launch.
``open_loop`` drives it with Poisson arrivals and reports latency percentiles.
"""
from __future__ import annotations

import ctypes
import random
import threading
import time
from typing import Dict, List, Sequence, Tuple

import numpy as np

from . import _native as N
from .api import _check


class Server:
    def __init__(self, hbm_budget_bytes: int, levels: int, device: int = 0):
        self._lib = N.load()
        h = ctypes.c_void_p()
        _check(self._lib.gpjson_server_create(device, int(hbm_budget_bytes), levels, ctypes.byref(h)))
        self._h = h
        self.levels = levels
        self.sizes: Dict[str, int] = {}

    def add(self, key: str, data) -> None:
        arr = np.frombuffer(memoryview(data), dtype=np.uint8) if not isinstance(data, np.ndarray) else \
            np.ascontiguousarray(data, dtype=np.uint8)
        _check(self._lib.gpjson_server_add(self._h, key.encode(), arr.ctypes.data_as(ctypes.c_void_p), arr.size))
        self.sizes[key] = int(arr.size)

    def admit(self, key: str) -> None:
        _check(self._lib.gpjson_server_admit(self._h, key.encode()))

    def submit(self, key: str, path: str) -> int:
        t = ctypes.c_uint64()
        _check(self._lib.gpjson_server_submit(self._h, key.encode(), path.encode(), ctypes.byref(t)))
        return t.value

    def ready(self, ticket: int, timeout_ms: int = -1) -> Tuple[int, int, float]:
        """Waits for a ticket without collecting it: (lines, results per line, latency us)."""
        lines, R, lat = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.gpjson_server_wait(self._h, ticket, timeout_ms, ctypes.byref(lines), ctypes.byref(R),
                                            ctypes.byref(lat), None, 0))
        return lines.value, R.value, lat.value

    def wait(self, ticket: int, timeout_ms: int = -1) -> Tuple[np.ndarray, float]:
        """(spans int64[lines][R][2], latency in microseconds) of a ticket."""
        lines, R, lat = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.gpjson_server_wait(self._h, ticket, timeout_ms, ctypes.byref(lines), ctypes.byref(R),
                                            ctypes.byref(lat), None, 0))
        out = np.empty((lines.value, R.value, 2), dtype=np.int64)
        _check(self._lib.gpjson_server_wait(self._h, ticket, timeout_ms, None, None, None,
                                            out.ctypes.data_as(ctypes.c_void_p), out.size))
        return out, lat.value

    def result(self, ticket: int, timeout_ms: int = -1) -> "ServedResult":
        """Zero-copy result: a view of the spans in the server's pinned memory,
        valid until ServedResult.release() (use it as a context manager)."""
        ptr = ctypes.POINTER(ctypes.c_int64)()
        lines, R, lat = ctypes.c_int64(), ctypes.c_int(), ctypes.c_double()
        _check(self._lib.gpjson_server_result(self._h, ticket, timeout_ms, ctypes.byref(ptr), ctypes.byref(lines),
                                              ctypes.byref(R), ctypes.byref(lat)))
        n = lines.value * R.value * 2
        view = np.ctypeslib.as_array(ptr, shape=(n,)).reshape(lines.value, R.value, 2) if n else \
            np.empty((lines.value, R.value, 2), dtype=np.int64)
        return ServedResult(self, ticket, view, lat.value)

    def stats(self) -> dict:
        v = (ctypes.c_int64 * 8)()
        _check(self._lib.gpjson_server_stats(self._h, v, 8))
        keys = ("resident_bytes", "resident_files", "launches", "requests", "max_batch", "admissions", "evictions",
                "spills")
        return dict(zip(keys, list(v)))

    def close(self) -> None:
        if self._h:
            self._lib.gpjson_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class ServedResult:
    """Spans of one served request, read in place (see Server.result)."""

    def __init__(self, server: Server, ticket: int, spans: np.ndarray, latency_us: float):
        self.server, self.ticket, self.spans, self.latency_us = server, ticket, spans, latency_us

    def release(self) -> None:
        if self.spans is not None:
            self.spans = None
            _check(self.server._lib.gpjson_server_release(self.server._h, self.ticket))

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.release()


def open_loop(server: Server, requests: Sequence[Tuple[str, str]], rate_per_s: float, seed: int = 0,
              keep_spans=False, wait_timeout_ms: int = 120000):
    """Submits `requests` with exponential inter-arrival times (Poisson arrivals
    at rate_per_s, seeded) from one thread while another collects results.
    Returns (latencies_us, wall_s, bytes_queried, spans or None); keep_spans
    is True (copy every request's spans) or a predicate on the request index.  A latency
    runs from the request's scheduled arrival, not from its submit call: when
    the submitting thread falls behind the schedule, that delay counts too
    (no coordinated omission).  The server measures submit to completion; the
    lag of the submit call behind the schedule is added to it."""
    keep = keep_spans if callable(keep_spans) else (lambda i: bool(keep_spans))
    rng = random.Random(seed)
    gaps = [rng.expovariate(rate_per_s) for _ in requests]
    tickets: List[int] = [0] * len(requests)
    ready = [threading.Event() for _ in requests]
    lat = [0.0] * len(requests)
    lag_us = [0.0] * len(requests)
    spans = [None] * len(requests)
    errors: List[BaseException] = []

    def collector():
        # results are read in place (zero copy) and released at once
        try:
            for i in range(len(requests)):
                ready[i].wait()
                with server.result(tickets[i], wait_timeout_ms) as res:
                    lat[i] = lag_us[i] + res.latency_us
                    if keep(i):
                        spans[i] = res.spans.copy()
        except BaseException as e:  # surfaced to the caller
            errors.append(e)

    th = threading.Thread(target=collector)
    th.start()
    t0 = time.perf_counter()
    due = t0
    for i, (key, path) in enumerate(requests):
        due += gaps[i]
        d = due - time.perf_counter()
        if d > 0:
            time.sleep(d)
        lag_us[i] = max(0.0, time.perf_counter() - due) * 1e6
        tickets[i] = server.submit(key, path)
        ready[i].set()
    th.join()
    wall = time.perf_counter() - t0
    if errors:
        raise errors[0]
    nbytes = sum(server.sizes[k] for k, _ in requests)
    return np.array(lat), wall, nbytes, (spans if keep_spans else None)
