"""GPU parity of the bytecode query kernel: every result span equals the
CPU oracle's restatement of the reference interpreter (query.cu:144-454)."""
import numpy as np
import pytest

import jsongen

pytestmark = pytest.mark.gpu

BOOKSTORE = (b'{"store":{"book":[{"category":"reference","author":"Nigel Rees","title":'
             b'"Sayings of the Century","price":"8.95"},{"category":"fiction","author":'
             b'"Herman Melville","title":"Moby Dick","isbn":"0-553-21311-3","price":"8.99"},'
             b'{"category":"fiction","author":"J.R.R. Tolkien","title":"The Lord of the Rings",'
             b'"isbn":"0-395-19395-8","price":"22.99"}],"bicycle":{"color":"red","price":'
             b'"19.95"}},"expensive":"10"}')

QUERIES = [
    "$.a", "$.b", "$.a.b", "$.a.c", "$.id", "$.name", "$.tags[0]", "$.tags[1]", "$.tags[-1]",
    "$.items[2]", "$.a[0:2]", "$.a[1,3]", "$.arr[:3]", "$.a[-2:]", "$.a[-1]", "$['a']['b']",
    "$.a[?(@ == \"x\")]", "$.h[?(@.k == \"v\")].id", "$.user.lang", "$.e[0].f", "$.x.y.a.b",
    "$.nested.arr[1].x", "$.k[0][1]", "$.a.b.c.d", "$", "$.ab", "$.abc",
]


def run_both(gpj, orc, data: bytes, query: str, levels=None):
    ir, depth, slots = gpj.compile_query(query)
    lv_count = max(depth, 1) if levels is None else levels
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpj.build_index(dev, len(data), lv_count)
    got = gpj.run_query(dev, len(data), idx, ir, slots)
    want = orc.query(data, orc.build_index(data, lv_count), ir, slots)
    return got, want


def texts(data: bytes, spans):
    out = []
    for line in spans:
        vals = [data[a:b] for a, b in line if a >= 0 and b >= 0]
        if vals:
            out.append(vals)
    return out


def test_bookstore_kats(gpu, orc):
    # tests/query/query_executor_test.cpp:155-245
    got, want = run_both(gpu, orc, BOOKSTORE, "$.expensive")
    np.testing.assert_array_equal(got, want)
    assert texts(BOOKSTORE, got) == [[b'"10"']]
    got, want = run_both(gpu, orc, BOOKSTORE, "$.store.book[1].author")
    np.testing.assert_array_equal(got, want)
    assert texts(BOOKSTORE, got) == [[b'"Herman Melville"']]
    got, want = run_both(gpu, orc, BOOKSTORE, "$.store.magazine")
    np.testing.assert_array_equal(got, want)
    assert texts(BOOKSTORE, got) == []


def test_engine_kats(gpu, orc):
    # tests/engine_test.cpp:95-128 (single partition form)
    data = b'{"name":"Ada","scores":[1,2],"lang":"en"}\n{"name":"Bob","scores":[3,4],"lang":"fr"}\n'
    got, want = run_both(gpu, orc, data, "$.name")
    np.testing.assert_array_equal(got, want)
    assert texts(data, got) == [[b'"Ada"'], [b'"Bob"']]
    got, want = run_both(gpu, orc, data, "$.scores[1]")
    np.testing.assert_array_equal(got, want)
    assert texts(data, got) == [[b"2"], [b"4"]]


@pytest.mark.parametrize("seed", [11, 12])
def test_random_records_all_queries(gpu, orc, seed):
    data = jsongen.ndjson(seed, 1 << 20, max_depth=6, spaces=(seed == 12))
    for q in QUERIES:
        print(f"[query] seed={seed} q={q}", flush=True)
        got, want = run_both(gpu, orc, data, q)
        if not np.array_equal(got, want):
            bad = np.argwhere((got != want).any(axis=(1, 2)))[:5].ravel().tolist()
            raise AssertionError(f"{q}: lines {bad} differ: got {got[bad].tolist()} "
                                 f"want {want[bad].tolist()}")


def test_extra_levels_do_not_change_results(gpu, orc):
    data = jsongen.ndjson(13, 200_000, max_depth=6)
    for q in ("$.a.b", "$.tags[1]"):
        base, _ = run_both(gpu, orc, data, q)
        more, want = run_both(gpu, orc, data, q, levels=12)
        np.testing.assert_array_equal(more, want)
        np.testing.assert_array_equal(base, more)


def test_byte_soup_queries_do_not_fault(gpu, orc):
    # arbitrary bytes: spans must still match the oracle's guarded semantics
    data = jsongen.byte_soup(21, 300_000, special_weight=0.3)
    for q in ("$.a", "$.a[2]", "$.a[-1]", "$.a.b[?(@ == \"x\")]"):
        got, want = run_both(gpu, orc, data, q)
        np.testing.assert_array_equal(got, want)


def test_key_lookup_forms(gpu, orc):
    # the byte before ':' decides how a key lookup finds the key's last byte:
    # an unescaped quote (compact keys), an escaped quote, a quote after an
    # escaped backslash, spaces, or no string at all
    lines = [
        b'{"a":1,"b":2,"ab":3}',
        b'{"k\\\\":1,"a":2}',
        b'{"a\\"":1,"a":2}',
        b'{x\\":1,"a":2}',
        b'{"a" :1,"b"  :2}',
        b'{"":1,"a":2}',
        b'{"b":{"a":[1,{"a":2}]},"a":3}',
        b'{ "a" : { "b" : 4 } }',
        b'{"\\\\":1,"a\\\\b":2}',
        b'{"x":"\\":","a":5}',
    ]
    data = b"\n".join(lines) + b"\n"
    for q in ("$.a", "$.b", "$.ab", "$.b.a", "$.a.b", "$['k\\\\']", "$.x"):
        print(f"[query] q={q}", flush=True)
        got, want = run_both(gpu, orc, data, q, levels=4)
        np.testing.assert_array_equal(got, want)


def test_key_search_ends_at_a_colon_level_end(gpu, orc):
    # malformed records whose level end is a ':' (see the oracle test of the
    # same name): the key search must end (the path fails) and match the
    # oracle; well-formed records around them are unaffected
    data = (b'{"a":{"x":1]:2}\n{"a":{"y":5},"b":1}\n{"a":{"x":1,"z":[3]]:7}\n'
            b'{"a":{"q":"w"}}\n') * 50
    for q in ("$.a.y", "$.a.z", "$.a.q"):
        print(f"[query] malformed level end q={q}", flush=True)
        got, want = run_both(gpu, orc, data, q)
        np.testing.assert_array_equal(got, want)


def batch_vs_oracle(gpj, orc, data: bytes, queries, levels: int, label: str):
    compiled = [gpj.compile_query(q, extensions=True) for q in queries]
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpj.build_index(dev, len(data), levels)
    oidx = orc.build_index(data, levels)
    batch = gpj.run_query_batch(dev, len(data), idx, compiled, capacity=4)
    for (ir, _d, slots), (spans, counts), q in zip(compiled, batch, queries):
        print(f"[batch] {label} q={q}", flush=True)
        if counts is None:
            np.testing.assert_array_equal(spans, orc.query(data, oidx, ir, slots))
        else:
            ws, wc = orc.query_ext(data, oidx, ir, spans.shape[1])
            np.testing.assert_array_equal(counts, wc)
            np.testing.assert_array_equal(spans, ws)


def test_batch_shared_first_key_edge_forms(gpu, orc):
    # a batch whose queries start with key lookups shares one scan of each
    # record's top-level object: every key form (compact, escaped, spaced,
    # long, empty, absent), malformed level ends and arbitrary bytes must give
    # each query's own results
    lines = [
        b'{"a":1,"b":2,"ab":3}',
        b'{"k\\\\":1,"a":2}',
        b'{"a\\"":1,"a":2}',
        b'{x\\":1,"a":2}',
        b'{"a" :1,"b"  :2}',
        b'{"":1,"a":2}',
        b'{"b":{"a":[1,{"a":2}]},"a":3}',
        b'{ "a" : { "b" : 4 } }',
        b'{"\\\\":1,"a\\\\b":2}',
        b'{"x":"\\":","a":5}',
        b'{"longerkey":{"b":1},"a":{"b":2},"longerkey2":7}',
        b'{"a":{"x":1]:2}',
        b'{"a":{"x":1,"z":[3]]:7}',
        b'{"b":1}}"a":2}',
        b'{"a":1,"b":2',
        b'{}',
        b'{"h":[{"t":"x","k":"v","id":1}],"a":[0,1]}',
    ]
    data = (b"\n".join(lines) + b"\n") * 3
    queries = ["$.a", "$.b", "$.ab", "$.b.a", "$['k\\\\']", "$.x", "$.longerkey.b"]
    batch_vs_oracle(gpu, orc, data, queries, 4, "forms")
    queries = ["$.a.b", "$.a[1]", "$.h[*].t", '$.h[?(@.k == "v")].id', "$.longerkey2", "$.a"]
    batch_vs_oracle(gpu, orc, data, queries, 4, "forms-ext")
    soup = jsongen.byte_soup(22, 300_000, special_weight=0.3)
    batch_vs_oracle(gpu, orc, soup, ["$.a", "$.a[2]", "$.b.a", '$.a.b[?(@ == "x")]'], 3, "soup")
    # one first-key query among others: the memo stays off
    batch_vs_oracle(gpu, orc, data, ["$.a", "$[0]", "$"], 4, "single")


def test_batch_shared_first_key_without_move_down(gpu, orc):
    # bytecode the compiler does not emit but the interpreter runs: MOVE_TO_KEY
    # "a" followed directly by STORE_RESULT + END, in a batch with `$.a`
    # (MOVE_TO_KEY "a", MOVE_DOWN, ...) and `$.b.a`: the shared search of "a"
    # must not take the MOVE_DOWN for both, and every result equals the
    # oracle's
    lines = [b'{"a":1,"b":{"a":2}}', b'{ "a" : [1, 2] ,"b":3}', b'{"b":{},"a":{"x":1}}',
             b'{"x":1}', b'{"a":1', b'{"b":{"a":"q"},"a" :  "s"}']
    data = (b"\n".join(lines) + b"\n") * 40
    raw = bytes([4, 1]) + b"a" + bytes([1, 0])  # MOVE_TO_KEY "a", STORE_RESULT, END
    compiled = [(raw, 1, 1), gpu.compile_query("$.a", extensions=True),
                gpu.compile_query("$.b.a", extensions=True)]
    dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpu.build_index(dev, len(data), 3)
    oidx = orc.build_index(data, 3)
    batch = gpu.run_query_batch(dev, len(data), idx, compiled, capacity=4)
    for (ir, _d, slots), (spans, counts) in zip(compiled, batch):
        print(f"[batch] raw-ir {ir!r}", flush=True)
        assert counts is None
        np.testing.assert_array_equal(spans, orc.query(data, oidx, ir, slots))
