"""Top-level array elements as records (extension): the oracle's element mode
pinned against two independent definitions -- line mode over the same bytes
with the top-level separators turned into newlines, and a DOM walk of the
parsed document."""
import json

import numpy as np
import pytest


def transformed(doc: bytes, bounds) -> bytes:
    """The document with its top-level ',' -> '\\n' and '[' ']' -> ' '."""
    d = bytearray(doc)
    for p in bounds:
        d[p] = ord("\n") if doc[p] == ord(",") else ord(" ")
    return bytes(d)


HAND_DOCS = [
    b'[{"id":1,"items":[{"sku":5,"name":"a"},{"sku":7,"name":"b\\"c"}]} , {"id":2,"items":[]},'
    b'{"x":{"items":[{"sku":9}]}},3,{"items":[{"sku":10}]}]',
    b"[]",
    b"[ ]",
    b"[1,2,3]",
    b'[{"items":[{"sku":"s,[x]"}]},[{"items":1}],{"items":[{"sku":{"a":[1,2]}},{"sku":2}]}]',
    b'  [ {"items" : [ {"sku" : 11 } , {"sku":12} ] } , {"items":[{"sku":13}]} ]  ',
    b'[{"items":[{"sku":1}]},{"items":[{"sku":2},{"sku":3}]}',  # unterminated
]
QUERIES = ["$.items[*].sku", "$.items[0].sku", "$.id", "$.items[*].name", "$.x.items[*].sku",
           '$.items[?(@.name == "a")].sku', "$.items", "$.meta"]


@pytest.mark.parametrize("doc", HAND_DOCS)
def test_element_mode_equals_line_mode_on_transformed_bytes(gpj, orc, doc):
    for q in QUERIES:
        ir, depth, slots = gpj.compile_query(q, extensions=True)
        variable = gpj.uses_extension_opcodes(ir)
        cap = 8 if variable else slots
        ix = orc.build_index(doc, depth + 1)
        b = orc.row_bounds(ix)
        spans, counts = orc.query_records(doc, ix, b, ir, cap, variable)
        t = transformed(doc, b)
        tix = orc.build_index(t, max(depth, 1))
        if b.size == 0:
            assert spans.shape[0] == 0
            continue
        if variable:
            ws, wc = orc.query_ext(t, tix, ir, cap)
            np.testing.assert_array_equal(counts, wc[: counts.size])
            np.testing.assert_array_equal(spans, ws[: spans.shape[0]])
        else:
            ws = orc.query(t, tix, ir, slots)
            np.testing.assert_array_equal(spans, ws[: spans.shape[0]])


def test_row_bounds_are_top_level_separators(orc):
    doc = HAND_DOCS[4]
    b = orc.row_bounds(orc.build_index(doc, 2))
    depth, in_str, esc, want = -1, False, False, []
    for i, ch in enumerate(doc):
        c = chr(ch)
        if in_str:
            if esc:
                esc = False
            elif c == "\\":
                esc = True
            elif c == '"':
                in_str = False
            continue
        if c == '"':
            in_str = True
        elif c in "[{":
            depth += 1
            if depth == 0:
                want.append(i)
        elif c in "]}":
            if depth == 0:
                want.append(i)
            depth -= 1
        elif c in ",:" and depth == 0:
            want.append(i)
    assert b.tolist() == want


def test_generated_array_matches_dom(gpj, orc):
    doc = bytes(gpj.generate(gpj.SCHEMA_BIG_ARRAY, 3000, 0x15418001))
    ir, depth, _ = gpj.compile_element_query("$[*].items[*].sku")
    ix = orc.build_index(doc, depth + 1)
    b = orc.row_bounds(ix)
    spans, counts = orc.query_records(doc, ix, b, ir, 32, True)
    elems = json.loads(doc)
    assert len(elems) == 3000 == counts.size
    for r, e in enumerate(elems):
        got = [int(doc[s:t]) for s, t in spans[r][: counts[r]]]
        assert got == [it["sku"] for it in e["items"]]
    assert counts.sum() > 0


def test_element_query_needs_leading_wildcard(gpj):
    with pytest.raises(gpj.MalformedQueryError):
        gpj.compile_element_query("$.items[*].sku")
    ir, depth, _ = gpj.compile_element_query("$[*].items[*].sku")
    assert ir == gpj.compile_query("$.items[*].sku", extensions=True)[0]
