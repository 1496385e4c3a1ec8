"""GPU parity of top-level-array element records (bounds from leveled row 0,
queries with interpreter level k on row k + 1) against the oracle."""
import numpy as np
import pytest

from test_array_records_cpu import HAND_DOCS, QUERIES

pytestmark = pytest.mark.gpu


def check_doc(gpj, orc, doc: bytes, queries, capacity=8):
    compiled = [gpj.compile_query(q, extensions=True) for q in queries]
    levels = max(c[1] for c in compiled) + 1
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(doc, dtype=np.uint8))
    idx = gpj.build_index(dev, len(doc), levels)
    bounds, m = gpj.array_bounds(idx)
    oix = orc.build_index(doc, levels)
    want_b = orc.row_bounds(oix)
    assert m == want_b.size
    if m:
        np.testing.assert_array_equal(bounds.to_host(np.int64, m), want_b)
    got = gpj.run_query_records(dev, len(doc), idx, bounds, m, compiled, capacity)
    for (ir, _d, slots), (spans, counts), q in zip(compiled, got, queries):
        print(f"[array] n={len(doc)} q={q}", flush=True)
        variable = counts is not None
        cap = capacity if variable else slots
        ws, wc = orc.query_records(doc, oix, want_b, ir, cap, variable)
        if variable:
            np.testing.assert_array_equal(counts, wc)
            ws = ws.copy()
            ws[np.arange(cap)[None, :] >= wc[:, None]] = -1
        np.testing.assert_array_equal(spans, ws)
    return got


@pytest.mark.parametrize("k", range(len(HAND_DOCS)))
def test_hand_documents(gpu, orc, k):
    check_doc(gpu, orc, HAND_DOCS[k], QUERIES)


def test_generated_array(gpu, orc):
    doc = bytes(gpu.generate(gpu.SCHEMA_BIG_ARRAY, 20000, 0x15418001))
    got = check_doc(gpu, orc, doc, ["$.items[*].sku", "$.id", "$.items[2].name", "$.meta"],
                    capacity=24)
    assert got[0][1].sum() > 0
