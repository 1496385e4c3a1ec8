"""GPU parity of the two query walks: one record per lane (64 records per
wave) and one record per wave over the record staged in LDS.  Launches of at
most 4,096 records take the wave walk by default, so most other GPU query
tests (small inputs) exercise it; here the main parity cases run
under each walk forced, plus what only the wave walk has: windows restaged
forwards (records longer than the 4 KiB window) and backwards (index searches
that leave the record), dense records beyond its 64 staged spans, array
element records, and inputs that are not 16-byte aligned (which take the lane
walk).  Every result must equal the oracle's."""
import numpy as np
import pytest

import jsongen
import test_gpu_query as tq
import test_gpu_query_ext as tx
from test_array_records_cpu import HAND_DOCS, QUERIES as ARRAY_QUERIES
from test_gpu_dense import check_dense
from test_gpu_wide_positions import mixed_data

pytestmark = pytest.mark.gpu

WALKS = {"lane": 1, "wave": 2}


@pytest.mark.parametrize("walk", sorted(WALKS))
@pytest.mark.parametrize("seed", [11, 12])
def test_reference_queries(gpu, orc, walk, seed):
    with gpu.walk(WALKS[walk]):
        tq.test_random_records_all_queries(gpu, orc, seed)


@pytest.mark.parametrize("walk", sorted(WALKS))
def test_kats_and_key_forms(gpu, orc, walk):
    with gpu.walk(WALKS[walk]):
        tq.test_bookstore_kats(gpu, orc)
        tq.test_engine_kats(gpu, orc)
        tq.test_key_lookup_forms(gpu, orc)
        tq.test_key_search_ends_at_a_colon_level_end(gpu, orc)
        tq.test_byte_soup_queries_do_not_fault(gpu, orc)


@pytest.mark.parametrize("walk", sorted(WALKS))
@pytest.mark.parametrize("seed", [31, 32])
def test_extension_queries(gpu, orc, walk, seed):
    with gpu.walk(WALKS[walk]):
        tx.test_extension_queries_match_oracle(gpu, orc, seed)


@pytest.mark.parametrize("walk", sorted(WALKS))
def test_batch_over_generated_records(gpu, orc, walk):
    with gpu.walk(WALKS[walk]):
        tx.test_batch_launch_equals_single_launches(gpu, orc)


def test_wave_windows_restaged_forwards(gpu, orc):
    # records up to ~13 KB: walks leave the first 4 KiB window
    data = mixed_data(44)
    assert max(len(x) for x in data.split(b"\n")) > 3 * 4096
    with gpu.walk(WALKS["wave"]):
        for q in tq.QUERIES:
            print(f"[walk] wave long records q={q}", flush=True)
            got, want = tq.run_both(gpu, orc, data, q)
            np.testing.assert_array_equal(got, want)
        for q in tx.EXT_QUERIES:
            print(f"[walk] wave long records ext q={q}", flush=True)
            spans, counts, ws, wc = tx.run_both_ext(gpu, orc, data, q, capacity=4)
            np.testing.assert_array_equal(counts, wc)
            np.testing.assert_array_equal(spans, ws)


def test_wave_windows_restaged_backwards(gpu, orc):
    # '{"a":]}' puts ']' on row 0, so $.a[-k] searches row 1 backwards out of
    # the record, across a 6 KB string, into earlier records; a key not
    # closed by a quote right before its ':' searches the string mask back
    # over 6 KB of spaces
    filler = b'{"s":"' + b"x" * 6000 + b'"}'
    lines = [b'{"a":[1,2,[3,4]],"b":{"c":[5]}}', filler, b'{"a":]}', b'{"a":[7, 8 ]}',
             filler, b'{"a": ,]}', b'{"a":[9]}', b'{"k":"' + b"\\\\" * 3000 + b'","a":1}',
             b'{"a":1,' + b" " * 6000 + b'x:2,"b":3}']
    data = b"\n".join(lines * 3) + b"\n"
    with gpu.walk(WALKS["wave"]):
        for q in ("$.a[-1]", "$.a[-2]", "$.a[-3:]", "$.a[0]", "$.b.c[-1]", "$.s", "$.a", "$.b"):
            print(f"[walk] wave backward q={q}", flush=True)
            got, want = tq.run_both(gpu, orc, data, q, levels=4)
            np.testing.assert_array_equal(got, want)


def test_wave_dense_records_beyond_the_stage(gpu, orc):
    # 0..150 values per record: records with more than 64 go to the fixup pass
    r = np.random.default_rng(8)
    lines = []
    for i in range(2000):
        k = int(r.integers(0, 151))
        vals = ",".join(str(int(v)) for v in r.integers(0, 1000, k))
        lines.append('{"id":%d,"a":[%s],"b":{"c":[%s]}}' % (i, vals, vals))
    data = ("\n".join(lines) + "\n").encode()
    with gpu.walk(WALKS["wave"]):
        check_dense(gpu, orc, data, ["$.a[*]", "$.b.c[*]", "$.id", "$.a[70]"], label="wave")


@pytest.mark.parametrize("walk", sorted(WALKS))
def test_array_element_records(gpu, orc, walk):
    compiled = [gpu.compile_query(q, extensions=True) for q in ARRAY_QUERIES]
    with gpu.walk(WALKS[walk]):
        for doc in HAND_DOCS:
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
                    ws, wc = orc.query_records(doc, oix, want_b, ir, 64 if variable else slots,
                                               variable)
                    wo, wsp = gpu.dense_from_slots(ws, wc if variable else None)
                    np.testing.assert_array_equal(offs, wo)
                    np.testing.assert_array_equal(spans, wsp)


def test_unaligned_input_takes_the_lane_walk(gpu, orc):
    data = jsongen.ndjson(61, 300_000, max_depth=5)
    n = len(data)
    host = np.zeros(n + 16, dtype=np.uint8)
    host[3:3 + n] = np.frombuffer(data, dtype=np.uint8)
    buf = gpu.DeviceBuffer.from_host(host)
    ptr = buf.ptr + 3
    for q in ("$.a.b", "$.tags[-1]", "$.user.lang"):
        ir, depth, slots = gpu.compile_query(q)
        levels = max(depth, 1)
        idx = gpu.build_index(ptr, n, levels)
        want = orc.query(data, orc.build_index(data, levels), ir, slots)
        with gpu.walk(WALKS["wave"]):
            got = gpu.run_query(ptr, n, idx, ir, slots)
        np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("schema", ["ndjson_1k", "ndjson_2k_4k"])
def test_walks_agree_on_generated_data(gpu, orc, schema):
    if schema == "ndjson_1k":
        data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_1K, 3 << 20, 0x15418002))
        queries = ["$.user.lang", "$.e[0].f", "$.h[*].t", '$.h[?(@.k == "v")].id']
    else:
        data = bytes(gpu.generate(gpu.SCHEMA_NDJSON_2_4K, 3 << 20, 0x15418000))
        queries = ["$.a.b", "$.a", "$.b"]
    compiled = [gpu.compile_query(q, extensions=True) for q in queries]
    levels = max(c[1] for c in compiled)
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpu.build_index(dev, len(data), levels)
    got = {}
    for walk in sorted(WALKS):
        with gpu.walk(WALKS[walk]):
            got[walk] = gpu.run_query_dense(dev, len(data), idx, compiled)
    oidx = orc.build_index(data, levels)
    for k, ((ir, _d, slots), q) in enumerate(zip(compiled, queries)):
        print(f"[walk] {schema} {q}", flush=True)
        if gpu.uses_extension_opcodes(ir):
            _, ct = orc.query_ext(data, oidx, ir, 1)
            cap = max(1, int(ct.max(initial=0)))
            ws, wc = orc.query_ext(data, oidx, ir, cap)
            wo, wsp = gpu.dense_from_slots(ws, wc)
        else:
            wo, wsp = gpu.dense_from_slots(orc.query(data, oidx, ir, slots))
        for walk in sorted(WALKS):
            np.testing.assert_array_equal(got[walk][k][0], wo)
            np.testing.assert_array_equal(got[walk][k][1], wsp)
