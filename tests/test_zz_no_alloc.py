"""GPU check of gpj_index_build_static, the index build that allocates
nothing (its scratch is a static device array of the library): on two calls
in a row, the second over a smaller input, every output equals
gpj_index_build_into's (caller-provided scratch) bit for bit, and so does a
query run over it."""
import ctypes

import numpy as np
import pytest

import jsongen

pytestmark = pytest.mark.gpu


def outputs(gpj, dev, n, levels, cap, scratch=None):
    """Builds with gpj_index_build_static (scratch None) or _into; returns
    host copies of (line count, newline entries, string mask, leveled rows)
    and the device buffers."""
    L = gpj.lib()
    w = (n + 63) // 64
    b = {"string": gpj.DeviceBuffer(w * 8), "leveled": gpj.DeviceBuffer(w * 8 * levels),
         "newline": gpj.DeviceBuffer(cap * 8), "lines": gpj.DeviceBuffer(8)}
    # stale contents must not leak into the results
    for k in ("string", "leveled", "newline"):
        gpj.check(L.gpj_h2d(b[k].ptr, np.full(b[k].nbytes, 0xA5, np.uint8).ctypes.data,
                            b[k].nbytes, None))
    if scratch is None:
        gpj.check(L.gpj_index_build_static(dev.ptr, n, levels, b["string"].ptr,
                                           b["leveled"].ptr, b["newline"].ptr, cap,
                                           b["lines"].ptr, None), "static build")
    else:
        gpj.check(L.gpj_index_build_into(dev.ptr, n, levels, b["string"].ptr,
                                         b["leveled"].ptr, b["newline"].ptr, cap,
                                         b["lines"].ptr, scratch.ptr, None), "build into")
    gpj.check(L.gpj_stream_sync(None))
    lines = int(b["lines"].to_host(np.int64, 1)[0])
    assert lines >= 1, f"failed build ({lines} lines)"
    host = (lines, b["newline"].to_host(np.int64, min(lines, cap)), b["string"].to_host(np.uint64, w),
            b["leveled"].to_host(np.uint64, w * levels))
    return host, b


def query(gpj, dev, n, b, levels, cap, lines):
    ir, depth, slots = gpj.compile_query("$.a.b")
    d_ir = gpj.DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8))
    spans = gpj.DeviceBuffer(lines * slots * 16)
    idx = gpj.IndexT(b["newline"].ptr, b["lines"].ptr, cap, b["string"].ptr, b["leveled"].ptr,
                     (n + 63) // 64, levels, 0)
    desc = gpj.QueryDescT(d_ir.ptr, len(ir), slots, 0, depth, spans.ptr, None)
    gpj.check(gpj.lib().gpj_query_run_batch(dev.ptr, n, ctypes.byref(idx), ctypes.byref(desc), 1,
                                            lines, None), "query")
    gpj.check(gpj.lib().gpj_stream_sync(None))
    return spans.to_host(np.int64, lines * slots * 2)


@pytest.mark.parametrize("order", ["static_first", "into_first"])
def test_static_scratch_equals_caller_scratch(gpu, order):
    levels = 3
    # the second input is smaller, so the static array keeps the first call's
    # larger descriptor area behind the second's
    for call, (seed, size) in enumerate([(31, 3 << 20), (32, 200_000)]):
        data = jsongen.ndjson(seed, size, max_depth=6)
        n = len(data)
        cap = n // 32 + 2
        print(f"[no-alloc] {order} call {call}: n={n} levels={levels}", flush=True)
        dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
        scratch = gpu.DeviceBuffer(gpu.lib().gpj_index_scratch_bytes(n))
        if order == "static_first":
            got, gb = outputs(gpu, dev, n, levels, cap)
            want, wb = outputs(gpu, dev, n, levels, cap, scratch)
        else:
            want, wb = outputs(gpu, dev, n, levels, cap, scratch)
            got, gb = outputs(gpu, dev, n, levels, cap)
        assert got[0] == want[0] == data.count(b"\n") + 1
        for g, w in zip(got[1:], want[1:]):
            np.testing.assert_array_equal(g, w)
        np.testing.assert_array_equal(query(gpu, dev, n, gb, levels, cap, got[0]),
                                      query(gpu, dev, n, wb, levels, cap, want[0]))


def test_static_build_rejects_inputs_above_its_limit(gpu):
    dummy = gpu.DeviceBuffer(64)
    st = gpu.lib().gpj_index_build_static(dummy.ptr, gpu.STATIC_MAX_BYTES + 1, 1, dummy.ptr,
                                          dummy.ptr, dummy.ptr, 1, dummy.ptr, None)
    assert st == gpu.ERR_CAPACITY
    assert b"static scratch" in gpu.lib().gpj_last_error()
