"""GPU checks of failure reporting and buffer bounds on the stream-ordered
paths.  A tile whose bounded wait for its carry gives up flags the build; the
tests force that from the host (a wait limit of 0 polls, so every tile that
would wait gives up at once and every kernel still ends normally) and check
that each caller reports it: the allocating build, the stream-ordered build
(line count -1, gpj_index_status), the multi-file batch and the server.  The
reference checks after every launch (src/gpjson/index/orig_index_builder.cu:107,115)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import jsongen

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "15418-project_amd", "_build", "gpjson_serve")


def ndjson(n_bytes: int, seed: int = 5) -> bytes:
    return jsongen.ndjson(seed, n_bytes, max_depth=4)


def build_into(gpj, dev, n, levels, cap, stream=None):
    """gpj_index_build_into with separately allocated buffers."""
    w = (n + 63) // 64
    bufs = {
        "string": gpj.DeviceBuffer(w * 8),
        "leveled": gpj.DeviceBuffer(w * 8 * levels),
        "newline": gpj.DeviceBuffer(cap * 8),
        "lines": gpj.DeviceBuffer(8),
        "scratch": gpj.DeviceBuffer(gpj.lib().gpj_index_scratch_bytes(n)),
    }
    gpj.check(gpj.lib().gpj_index_build_into(dev.ptr, n, levels, bufs["string"].ptr,
                                             bufs["leveled"].ptr, bufs["newline"].ptr, cap,
                                             bufs["lines"].ptr, bufs["scratch"].ptr, stream),
              "index build into")
    gpj.check(gpj.lib().gpj_stream_sync(stream))
    return bufs


def test_allocating_build_reports_failure(gpu):
    data = ndjson(1 << 20)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    print(f"[status] allocating build n={len(data)} wait_limit=0", flush=True)
    with gpu.wait_limit(0):
        with pytest.raises(gpu.DeviceError, match="timed out"):
            gpu.build_index(dev, len(data), 3)
    # the default limit builds the same input
    idx = gpu.build_index(dev, len(data), 3)
    assert idx.num_lines == data.count(b"\n") + 1


def test_stream_ordered_build_reports_failure(gpu, orc):
    data = ndjson(1 << 20, seed=6)
    n = len(data)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    cap = n // 64 + 2
    print(f"[status] build_into n={n} wait_limit=0", flush=True)
    with gpu.wait_limit(0):
        bufs = build_into(gpu, dev, n, 2, cap)
    assert int(bufs["lines"].to_host(np.int64, 1)[0]) == -1
    with pytest.raises(gpu.DeviceError, match="timed out"):
        gpu.index_status(bufs["scratch"], n)
    # a query over the failed index runs no line: the counts keep their fill
    ir, depth, slots = gpu.compile_query("$.a.b")
    d_ir = gpu.DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8))
    spans = gpu.DeviceBuffer.from_host(np.full(cap * slots * 2, 7, dtype=np.int64))
    counts = gpu.DeviceBuffer.from_host(np.full(cap, 7, dtype=np.int64))
    idx = gpu.IndexT(bufs["newline"].ptr, bufs["lines"].ptr, cap, bufs["string"].ptr,
                     bufs["leveled"].ptr, (n + 63) // 64, 2, 0)
    desc = gpu.QueryDescT(d_ir.ptr, len(ir), slots, 0, depth, spans.ptr, counts.ptr)
    gpu.check(gpu.lib().gpj_query_run_batch(dev.ptr, n, ctypes.byref(idx), ctypes.byref(desc),
                                            1, cap, None), "query")
    gpu.check(gpu.lib().gpj_stream_sync(None))
    assert (counts.to_host(np.int64, cap) == 7).all()
    # the same buffers rebuilt with the default limit: status clear, exact index
    bufs2 = build_into(gpu, dev, n, 2, cap)
    gpu.index_status(bufs2["scratch"], n)
    s, lv, nl = orc.build_index(data, 2)
    assert int(bufs2["lines"].to_host(np.int64, 1)[0]) == nl.size
    np.testing.assert_array_equal(bufs2["newline"].to_host(np.int64, nl.size), nl)


def test_batch_reports_failure(gpu):
    from batchdata import BATCH_QUERIES, plan
    schemas, targets, masks = plan(12, 1 << 16, 0.5, 1 << 14, 1 << 18, 0x5151)
    buf, off, sz = gpu.generate_files(schemas, targets, 0x5151)
    compiled = [gpu.compile_query(q, extensions=True) for q in BATCH_QUERIES]
    b = gpu.FileBatch(buf, off, sz, masks, compiled, capacity=8, stream=None)
    print(f"[status] batch of {b.num_files} files, {b.n} bytes, wait_limit=0", flush=True)
    with gpu.wait_limit(0):
        b.launch()
        with pytest.raises(gpu.DeviceError, match="timed out"):
            b.device_line_prefix()
        with pytest.raises(gpu.DeviceError, match="timed out"):
            b.fetch()
        b.per_file_setup()
        b.launch_per_file()
        with pytest.raises(gpu.DeviceError, match="timed out"):
            b.fetch(per_file=True)
    b.launch()
    np.testing.assert_array_equal(b.device_line_prefix(), b.line_prefix)


def test_server_reports_failure():
    args = [SERVE, "--requests", "6", "--rates", "1000", "--min-doc", "256KiB",
            "--max-doc", "512KiB", "--slots", "2", "--verify", "0", "--seed", "21",
            "--debug-wait-limit", "0"]
    print(f"[status] {' '.join(args[1:])}", flush=True)
    p = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert p.returncode == 2, (p.returncode, p.stdout[-500:], p.stderr[-1000:])
    assert "timed out" in p.stderr


def test_newline_count_above_capacity_is_clamped(gpu, orc):
    # more lines than the newline buffer holds: the query reads no entry past
    # the buffer (the count is clamped to the capacity) and the lines that fit
    # whole still match the oracle
    data = ndjson(300000, seed=9)
    n = len(data)
    s, lv, nl = orc.build_index(data, 2)
    cap = nl.size // 2
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    print(f"[status] n={n} lines={nl.size} newline capacity={cap}", flush=True)
    bufs = build_into(gpu, dev, n, 2, cap)
    assert int(bufs["lines"].to_host(np.int64, 1)[0]) == nl.size
    np.testing.assert_array_equal(bufs["newline"].to_host(np.int64, cap), nl[:cap])
    ir, depth, slots = gpu.compile_query("$.a.b")
    d_ir = gpu.DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8))
    spans = gpu.DeviceBuffer(cap * slots * 16)
    idx = gpu.IndexT(bufs["newline"].ptr, bufs["lines"].ptr, cap, bufs["string"].ptr,
                     bufs["leveled"].ptr, (n + 63) // 64, 2, 0)
    desc = gpu.QueryDescT(d_ir.ptr, len(ir), slots, 0, depth, spans.ptr, None)
    gpu.check(gpu.lib().gpj_query_run_batch(dev.ptr, n, ctypes.byref(idx), ctypes.byref(desc),
                                            1, cap, None), "query")
    gpu.check(gpu.lib().gpj_stream_sync(None))
    got = spans.to_host(np.int64, cap * slots * 2).reshape(cap, slots, 2)
    want = orc.query(data, (s, lv, nl), ir, slots)
    np.testing.assert_array_equal(got[:cap - 1], want[:cap - 1])
