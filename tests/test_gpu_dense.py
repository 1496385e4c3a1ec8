"""GPU parity of dense results (gpj_query_run_dense and
gpj_query_run_records_dense: counts, scan and emission on the device, one
span array per query in record order plus CSR offsets) against the oracle's
per-record results put in the same form.  Covers fixed-slot and extension
bytecode in one launch, records with more spans than the LDS stage holds (the
fixup pass), outputs too small for the total, array-element records, the
64-bit position pass and a failed index build."""

import numpy as np
import pytest

import jsongen
from test_array_records_cpu import HAND_DOCS, QUERIES as ARRAY_QUERIES
from test_gpu_query import QUERIES as REF_QUERIES
from test_gpu_query_ext import EXT_QUERIES
from test_gpu_wide_positions import SPAN, mixed_data

pytestmark = pytest.mark.gpu


def oracle_dense(gpj, orc, data, oidx, ir, slots):
    """The oracle's results of one query as (offsets, spans [total, 2])."""
    if gpj.uses_extension_opcodes(ir):
        _, counts = orc.query_ext(data, oidx, ir, 1)
        cap = max(1, int(counts.max(initial=0)))
        spans, counts = orc.query_ext(data, oidx, ir, cap)
        return gpj.dense_from_slots(spans, counts)
    return gpj.dense_from_slots(orc.query(data, oidx, ir, slots))


def check_dense(gpj, orc, data: bytes, queries, capacity=None, label=""):
    compiled = [gpj.compile_query(q, extensions=True) for q in queries]
    levels = max(1, max(c[1] for c in compiled))
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpj.build_index(dev, len(data), levels)
    oidx = orc.build_index(data, levels)
    got = gpj.run_query_dense(dev, len(data), idx, compiled, capacity=capacity)
    for (ir, _d, slots), (offs, spans), q in zip(compiled, got, queries):
        print(f"[dense] {label} n={len(data)} q={q}", flush=True)
        wo, ws = oracle_dense(gpj, orc, data, oidx, ir, slots)
        np.testing.assert_array_equal(offs, wo)
        np.testing.assert_array_equal(spans, ws)
    return got


@pytest.mark.parametrize("seed", [51, 52])
def test_dense_random_records(gpu, orc, seed):
    data = jsongen.ndjson(seed, 1 << 20, max_depth=6, spaces=(seed == 52))
    # at most 8 queries per launch
    qs = EXT_QUERIES + REF_QUERIES
    for k in range(0, len(qs), 8):
        check_dense(gpu, orc, data, qs[k:k + 8], label=f"seed={seed}")


def test_dense_throughput_records_one_launch(gpu, orc):
    data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 4 << 20, 0x15418002))
    got = check_dense(gpu, orc, data,
                      ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id'],
                      label="throughput")
    assert all(int(o[-1]) > 0 for o, _ in got)


def test_dense_records_beyond_the_stage(gpu, orc):
    # 0..40 values per record: records with more than 8 go through the fixup
    r = np.random.default_rng(5)
    lines = []
    for i in range(3000):
        k = int(r.integers(0, 41))
        vals = ",".join(str(int(v)) for v in r.integers(0, 1000, k))
        lines.append('{"id":%d,"a":[%s],"b":{"c":[%s]}}' % (i, vals, vals))
    data = ("\n".join(lines) + "\n").encode()
    with gpu.walk(gpu.WALK_LANE):  # (the lane walk's stage holds 8 spans per record)
        check_dense(gpu, orc, data, ["$.a[*]", "$.b.c[*]", "$.id", "$.a[3]"], label="stage")


def test_dense_output_smaller_than_total(gpu, orc):
    data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 1 << 20, 0x15418002))
    ir, depth, slots = gpu.compile_query("$.h[*].t", extensions=True)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpu.build_index(dev, len(data), depth)
    wo, ws = oracle_dense(gpu, orc, data, orc.build_index(data, depth), ir, slots)
    total = int(wo[-1])
    (offs, spans), = gpu.run_query_dense(dev, len(data), idx, [(ir, depth, slots)],
                                         capacity=total // 3, rerun=False)
    np.testing.assert_array_equal(offs, wo)  # the offsets are complete
    np.testing.assert_array_equal(spans, ws[:total // 3])
    (offs, spans), = gpu.run_query_dense(dev, len(data), idx, [(ir, depth, slots)],
                                         capacity=total // 3)  # reruns with the total
    np.testing.assert_array_equal(spans, ws)


def test_dense_no_match_anywhere(gpu, orc):
    # (records without an empty object: MOVE_TO_KEY falls through `{}`)
    data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 300000, 0x15418002))
    got = check_dense(gpu, orc, data, ["$.nosuchkey", "$.nosuchkey[*].x"], label="none")
    assert all(int(o[-1]) == 0 for o, _ in got)


@pytest.mark.parametrize("k", range(len(HAND_DOCS)))
def test_dense_array_elements(gpu, orc, k):
    doc = HAND_DOCS[k]
    compiled = [gpu.compile_query(q, extensions=True) for q in ARRAY_QUERIES]
    levels = max(c[1] for c in compiled) + 1
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(doc, dtype=np.uint8))
    idx = gpu.build_index(dev, len(doc), levels)
    bounds, m = gpu.array_bounds(idx)
    oix = orc.build_index(doc, levels)
    want_b = orc.row_bounds(oix)
    for start in range(0, len(compiled), 8):
        part = compiled[start:start + 8]
        got = gpu.run_query_dense(dev, len(doc), idx, part, bounds=bounds, num_bounds=m)
        for (ir, _d, slots), (offs, spans) in zip(part, got):
            variable = gpu.uses_extension_opcodes(ir)
            cap = 64 if variable else slots
            ws, wc = orc.query_records(doc, oix, want_b, ir, cap, variable)
            wo, wsp = gpu.dense_from_slots(ws, wc if variable else None)
            np.testing.assert_array_equal(offs, wo)
            np.testing.assert_array_equal(spans, wsp)


def test_dense_generated_array(gpu, orc):
    doc = bytes(gpu.generate(gpu.SCHEMA_BIG_ARRAY, 20000, 0x15418001))
    compiled = [gpu.compile_element_query(q) for q in ("$[*].items[*].sku", "$[*].id")]
    levels = max(c[1] for c in compiled) + 1
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(doc, dtype=np.uint8))
    idx = gpu.build_index(dev, len(doc), levels)
    bounds, m = gpu.array_bounds(idx)
    oix = orc.build_index(doc, levels)
    want_b = orc.row_bounds(oix)
    got = gpu.run_query_dense(dev, len(doc), idx, compiled, bounds=bounds, num_bounds=m)
    for (ir, _d, slots), (offs, spans) in zip(compiled, got):
        variable = gpu.uses_extension_opcodes(ir)
        ws, wc = orc.query_records(doc, oix, want_b, ir, 32 if variable else slots, variable)
        wo, wsp = gpu.dense_from_slots(ws, wc if variable else None)
        np.testing.assert_array_equal(offs, wo)
        np.testing.assert_array_equal(spans, wsp)
    assert int(got[0][0][-1]) > 8 * (m - 1)  # most elements hold several items


def test_dense_wide_positions(gpu, orc):
    data = mixed_data(43)
    qs = ["$.a.b[*].x", "$.a.b[149].s", "$.h[*].t", "$.tags[*]", "$.user.lang", "$.a.b[-1].x"]
    want = check_dense(gpu, orc, data, qs, label="default span")
    with gpu.narrow_span(SPAN), gpu.walk(gpu.WALK_LANE):
        got = check_dense(gpu, orc, data, qs, label=f"span={SPAN}")
    for (a, b), (c, d) in zip(want, got):
        np.testing.assert_array_equal(a, c)
        np.testing.assert_array_equal(b, d)


def test_dense_after_failed_build_has_no_records(gpu):
    data = jsongen.ndjson(8, 1 << 20, max_depth=4)
    n = len(data)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    cap, w, levels = n // 64 + 2, (n + 63) // 64, 2
    bufs = [gpu.DeviceBuffer(w * 8), gpu.DeviceBuffer(w * 8 * levels), gpu.DeviceBuffer(cap * 8),
            gpu.DeviceBuffer(8), gpu.DeviceBuffer(gpu.lib().gpj_index_scratch_bytes(n))]
    with gpu.wait_limit(0):
        gpu.check(gpu.lib().gpj_index_build_into(dev.ptr, n, levels, bufs[0].ptr, bufs[1].ptr,
                                                 bufs[2].ptr, cap, bufs[3].ptr, bufs[4].ptr,
                                                 None), "build")
    raw = gpu.IndexT(bufs[2].ptr, bufs[3].ptr, cap, bufs[0].ptr, bufs[1].ptr, w, levels, 0)
    idx = gpu.BuiltIndex(raw, n)
    compiled = [gpu.compile_query(q, extensions=True) for q in ("$.a", "$.h[*].t")]
    got = gpu.run_query_dense(dev, n, idx, compiled, capacity=16, max_records=cap)
    for offs, spans in got:
        assert (offs == 0).all() and spans.shape == (0, 2)
