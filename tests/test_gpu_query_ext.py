"""GPU parity of the query extensions ([*] wildcard, array-element filter,
variable-length results) against the oracle's restatement of the same
semantics."""
import numpy as np
import pytest

import jsongen

pytestmark = pytest.mark.gpu

EXT_QUERIES = ["$.h[*].t", '$.h[?(@.k == "v")].id', "$.tags[*]", "$.a[*].b", "$.a[*][*]",
               "$.arr[*].x.y", '$.a[?(@ == "x")]', "$.items[*].sku", "$.e[0].f", "$.user.lang"]


def run_both_ext(gpj, orc, data: bytes, query: str, capacity: int):
    ir, depth, _ = gpj.compile_query(query, extensions=True)
    levels = max(depth, 1)
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpj.build_index(dev, len(data), levels)
    spans, counts = gpj.run_query_ext(dev, len(data), idx, ir, capacity)
    cap = spans.shape[1]
    want_spans, want_counts = orc.query_ext(data, orc.build_index(data, levels), ir, cap)
    return spans, counts, want_spans, want_counts


@pytest.mark.parametrize("seed", [31, 32])
def test_extension_queries_match_oracle(gpu, orc, seed):
    data = jsongen.ndjson(seed, 1 << 20, max_depth=6, spaces=(seed == 32))
    for q in EXT_QUERIES:
        print(f"[ext] seed={seed} q={q}", flush=True)
        spans, counts, ws, wc = run_both_ext(gpu, orc, data, q, capacity=4)
        np.testing.assert_array_equal(counts, wc)
        np.testing.assert_array_equal(spans, ws)


def test_extension_on_generated_throughput_records(gpu, orc):
    data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 4 << 20, 0x15418002))
    for q in ["$.h[*].t", '$.h[?(@.k == "v")].id']:
        spans, counts, ws, wc = run_both_ext(gpu, orc, data, q, capacity=8)
        np.testing.assert_array_equal(counts, wc)
        np.testing.assert_array_equal(spans, ws)
        assert counts.sum() > 0


def test_batch_launch_equals_single_launches(gpu, orc):
    data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 2 << 20, 0x15418002))
    queries = ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id', "$.arr[0:2]",
               "$.h[1].k"]
    compiled = [gpu.compile_query(q, extensions=True) for q in queries]
    levels = max(c[1] for c in compiled)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpu.build_index(dev, len(data), levels)
    batch = gpu.run_query_batch(dev, len(data), idx, compiled, capacity=8)
    oidx = orc.build_index(data, levels)
    for (ir, _d, slots), (spans, counts), q in zip(compiled, batch, queries):
        print(f"[batch] {q}", flush=True)
        if counts is None:
            np.testing.assert_array_equal(spans, gpu.run_query(dev, len(data), idx, ir, slots))
            np.testing.assert_array_equal(spans, orc.query(data, oidx, ir, slots))
        else:
            ws, wc = orc.query_ext(data, oidx, ir, spans.shape[1])
            np.testing.assert_array_equal(counts, wc)
            np.testing.assert_array_equal(spans, ws)
