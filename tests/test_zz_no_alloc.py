"""gpjson_query_noalloc (the plain path with every scratch array in one fixed
device array, for callers that cannot allocate) against gpjson_plan_create +
gpjson_plan_run, which run the same kernels with library-allocated scratch:
newline index and spans must be equal bit for bit, on two calls in a row, the
second with a smaller input (the fixed array keeps the first call's contents).
Inputs and outputs are caller-owned device buffers, allocated here through
the HIP runtime the library itself links (libamdhip64.so.7), since the
caller, not the library, owns them.
"""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


class DeviceArray:
    """Caller-owned device memory holding a numpy array's bytes."""

    def __init__(self, arr: np.ndarray):
        self.hip = ctypes.CDLL("libamdhip64.so.7")
        self.arr = np.ascontiguousarray(arr)
        self.nbytes = max(self.arr.nbytes, 8)
        self.ptr = ctypes.c_void_p()
        assert self.hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.nbytes)) == 0
        if self.arr.nbytes:
            assert self.hip.hipMemcpy(self.ptr, self.arr.ctypes.data_as(ctypes.c_void_p),
                                      ctypes.c_size_t(self.arr.nbytes), 1) == 0  # host to device

    def read(self) -> np.ndarray:
        out = np.empty_like(self.arr)
        if out.nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                      ctypes.c_size_t(out.nbytes), 2) == 0  # device to host
        return out

    def free(self) -> None:
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = ctypes.c_void_p()


def _noalloc(lib, ctx, data, irs, newline_cap, span_cap):
    inp = DeviceArray(np.frombuffer(data, dtype=np.uint8))
    nl = DeviceArray(np.full(newline_cap, -7, dtype=np.int64))
    sp = DeviceArray(np.full(max(span_cap, 1), -7, dtype=np.int64))
    info = DeviceArray(np.full(2, -7, dtype=np.int64))
    nq = len(irs)
    bufs = [ctypes.create_string_buffer(bytes(ir), max(len(ir), 1)) for ir in irs]
    ptrs = (ctypes.c_void_p * nq)(*[ctypes.cast(b, ctypes.c_void_p) for b in bufs])
    lens = (ctypes.c_int64 * nq)(*[len(ir) for ir in irs])
    rc = lib.gpjson_query_noalloc(ctx.handle, inp.ptr, len(data), ptrs, lens, nq, nl.ptr, newline_cap, sp.ptr,
                                  span_cap, info.ptr)
    ctx.synchronize()
    out = rc, nl.read(), sp.read(), info.read()
    for a in (inp, nl, sp, info):
        a.free()
    return out


def _reference(ctx, data, cps):
    """The existing entry point: a plan (library-allocated scratch), run once."""
    buf = ctx.stage(data)
    plan = ctx.plan(buf, [c.ir for c in cps])
    plan.run()
    spans = plan.spans()
    idx = ctx.index(buf, max(c.max_depth for c in cps))
    nl = idx.newline_index()
    idx.free()
    plan.free()
    buf.free()
    return nl, spans


def test_noalloc_equals_plan_two_calls(ctx):
    from gpjson_amd import _native as N
    from gpjson_amd import compile_path, datagen
    lib = N.load()
    cases = [
        ("tweets 12 MiB, 3 paths", datagen.generate("tweets", 501, 12 << 20).tobytes(),
         ["$.user.lang", "$.place.country_code", "$.user.stats.followers"]),
        ("adversarial 3 MiB, 2 paths", datagen.generate("adversarial", 502, 3 << 20).tobytes(),
         ["$.arr[999]", "$.chain.c1.c2"]),
    ]
    for name, data, paths in cases:
        cps = [compile_path(p) for p in paths]
        R = max(c.num_results for c in cps)
        want_nl, want_spans = _reference(ctx, data, cps)
        lines = want_nl.size
        print(f"[noalloc] {name}: {len(data)} bytes, {lines} lines", flush=True)
        cap_nl = lines + 100
        cap_sp = len(cps) * cap_nl * R * 2
        rc, nl, sp, info = _noalloc(lib, ctx, data, [c.ir for c in cps], cap_nl, cap_sp)
        assert rc == 0, N.last_error()
        assert info[0] == lines and info[1] == 0, (name, info)
        assert np.array_equal(nl[:lines], want_nl), name
        got = sp[:len(cps) * lines * R * 2].reshape(len(cps), lines, R, 2)
        assert np.array_equal(got, want_spans), name


def test_noalloc_limits(ctx):
    from gpjson_amd import _native as N
    from gpjson_amd import compile_path, datagen
    lib = N.load()
    cp = compile_path("$.user.lang")
    # an input whose index needs more than the fixed array: rejected before any launch
    big = 64 << 20
    inp = DeviceArray(np.zeros(big, dtype=np.uint8))
    out = DeviceArray(np.zeros(16, dtype=np.int64))
    ir = ctypes.create_string_buffer(cp.ir, len(cp.ir))
    ptrs = (ctypes.c_void_p * 1)(ctypes.cast(ir, ctypes.c_void_p))
    lens = (ctypes.c_int64 * 1)(len(cp.ir))
    rc = lib.gpjson_query_noalloc(ctx.handle, inp.ptr, big, ptrs, lens, 1, out.ptr, 4, out.ptr, 8, out.ptr)
    inp.free()
    out.free()
    assert rc == N.GPJSON_ERR_CAPACITY, rc
    assert "scratch" in N.last_error()
    # more lines than the caller's buffers hold: info[1] = 1 and no query runs
    data = datagen.generate("tweets", 503, 1 << 20).tobytes()
    rc, nl, sp, info = _noalloc(lib, ctx, data, [cp.ir], 10, 10 * 2)
    assert rc == 0, N.last_error()
    assert info[1] == 1 and info[0] > 10
    assert np.all(sp == -7), "spans were written although the line count exceeded their capacity"
