"""GPU parity of the query kernels' 64-bit pass.  The narrow query kernels
hold positions as 32-bit offsets from each record's start and hand a record
to the 64-bit kernel when it reaches the span from its start or a backward
search reaches the span below it.  With the span lowered from 2^31 - 2^20 to
4 KiB both cases occur on small inputs, and every result must still equal
the oracle's and the default run's."""
import numpy as np
import pytest

import jsongen
from batchdata import BATCH_QUERIES, plan
from test_gpu_query import QUERIES, run_both
from test_gpu_query_ext import EXT_QUERIES, run_both_ext

pytestmark = pytest.mark.gpu
SPAN = 4096


def long_record(i: int) -> bytes:
    items = b",".join(b'{"x":%d,"s":"%s","k":"v"}' % (k, b"y" * (8 + k % 50))
                      for k in range(150 + 7 * i))
    return (b'{"a":{"b":[' + items + b'],"c":"w"},"tags":[1,[2,3],4],"name":"L%d",'
            b'"h":[{"t":"p","k":"v","id":%d},{"t":"q","k":"u","id":9}],"user":{"lang":"en"},'
            b'"e":[{"f":1},{"f":2}],"items":[{"sku":5},{"sku":6}],"arr":[{"x":{"y":1}}]}'
            % (i, i))


def mixed_data(seed: int) -> bytes:
    lines = jsongen.ndjson(seed, 1 << 19, max_depth=6).split(b"\n")
    out = []
    for k, line in enumerate(lines):
        out.append(line)
        if k % 40 == 7:
            out.append(long_record(k // 40 % 13))
    return b"\n".join(out)


def test_long_records_fixed_queries(gpu, orc):
    data = mixed_data(41)
    assert max(len(x) for x in data.split(b"\n")) > 2 * SPAN
    for q in QUERIES:
        print(f"[wide] fixed q={q}", flush=True)
        with gpu.narrow_span(SPAN), gpu.walk(gpu.WALK_LANE):
            got, want = run_both(gpu, orc, data, q)
        np.testing.assert_array_equal(got, want)


def test_long_records_extension_queries(gpu, orc):
    data = mixed_data(42)
    for q in EXT_QUERIES:
        print(f"[wide] ext q={q}", flush=True)
        with gpu.narrow_span(SPAN), gpu.walk(gpu.WALK_LANE):
            spans, counts, ws, wc = run_both_ext(gpu, orc, data, q, capacity=4)
        np.testing.assert_array_equal(counts, wc)
        np.testing.assert_array_equal(spans, ws)


def test_backward_search_past_the_span(gpu, orc):
    # '{"a":]}' puts ']' on row 0, so $.a[-k] searches row 1 backwards out of
    # the record, across a 6 KB string, into earlier records
    filler = b'{"s":"' + b"x" * 6000 + b'"}'
    lines = [b'{"a":[1,2,[3,4]],"b":{"c":[5]}}', filler, b'{"a":]}', b'{"a":[7, 8 ]}',
             filler, b'{"a": ,]}', b'{"a":[9]}']
    data = b"\n".join(lines * 3) + b"\n"
    for q in ("$.a[-1]", "$.a[-2]", "$.a[-3:]", "$.a[0]", "$.b.c[-1]", "$.s"):
        print(f"[wide] backward q={q}", flush=True)
        with gpu.narrow_span(SPAN), gpu.walk(gpu.WALK_LANE):
            got, want = run_both(gpu, orc, data, q, levels=4)
        np.testing.assert_array_equal(got, want)
        base, _ = run_both(gpu, orc, data, q, levels=4)
        np.testing.assert_array_equal(got, base)


def test_batch_marks_in_both_result_modes(gpu, orc):
    # the mark sits in the first query's results: a fixed-slot query first,
    # then a variable-length one first
    data = mixed_data(43)
    for queries in (["$.name", "$.h[*].t", "$.a.c"], ["$.h[*].t", "$.name", "$.tags[-1]"]):
        compiled = [gpu.compile_query(q, extensions=True) for q in queries]
        levels = max(c[1] for c in compiled)
        dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
        idx = gpu.build_index(dev, len(data), levels)
        with gpu.narrow_span(SPAN), gpu.walk(gpu.WALK_LANE):
            batch = gpu.run_query_batch(dev, len(data), idx, compiled, capacity=8)
        oidx = orc.build_index(data, levels)
        for (ir, _d, slots), (spans, counts), q in zip(compiled, batch, queries):
            print(f"[wide] batch {q}", flush=True)
            if counts is None:
                np.testing.assert_array_equal(spans, orc.query(data, oidx, ir, slots))
            else:
                ws, wc = orc.query_ext(data, oidx, ir, spans.shape[1])
                np.testing.assert_array_equal(counts, wc)
                np.testing.assert_array_equal(spans, ws)


def test_file_batch_with_marks_equals_default(gpu):
    schemas, targets, masks = plan(24, 12288, 1.5, 1024, 1 << 17, 0x15418013)
    buf, off, sz = gpu.generate_files(schemas, targets, 0x15418013)
    compiled = [gpu.compile_query(q, extensions=True) for q in BATCH_QUERIES]
    b = gpu.FileBatch(buf, off, sz, masks, compiled, capacity=8, stream=None)
    b.launch()
    want_spans, want_counts = b.fetch()
    with gpu.narrow_span(SPAN):
        b.launch()
        spans, counts = b.fetch()
    for k in range(2):
        np.testing.assert_array_equal(counts[k], want_counts[k])
        np.testing.assert_array_equal(spans[k], want_spans[k])
