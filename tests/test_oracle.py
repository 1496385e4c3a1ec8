"""Pins the CPU oracle (oracle/gpj_oracle.c) against the reference's own
known-answer tests and against independent computations, so a plausible
mistake in the restatement (dropped term, wrong sign, off-by-one) fails here.
CPU only."""
import itertools
import random

import numpy as np
import pytest

import jsongen


def bits(words, n):
    return [int((int(words[i // 64]) >> (i % 64)) & 1) for i in range(n)]


# ---------------------------------------------------------------------------
# reference KATs
# ---------------------------------------------------------------------------
def test_kat_string_index(orc):
    # tests/index/string_index_test.cpp:86-93
    assert int(orc.string_index(b'"a\\"b"  "x"')[0]) == 0b01100011111


def test_kat_newline_index(orc):
    # tests/index/newline_index_test.cpp:64-69
    assert orc.newline_index(b"first\nsecond\nthird").tolist() == [0, 5, 12]


def test_kat_leveled_index(orc):
    # tests/index/leveled_bitmap_index_test.cpp:123-131
    lv = orc.leveled_index(b'{"a":[1,2]}', 3)
    assert [int(x) for x in lv[:, 0]] == [0b10000010001, 0b1010100000, 0]


def test_kat_partitions(orc):
    # tests/file/file_reader_test.cpp:89-155
    assert orc.partitions(b"alpha\nbeta\ngamma", 7) == [(0, 5), (6, 10), (11, 16)]
    data = b'{"id":1}\n{"id":2}\n'
    assert orc.partitions(data, 0) == [(0, len(data))]
    assert orc.partitions(b"", 64) == [(0, 0)]
    with pytest.raises(orc.PartitionError):
        orc.partitions(b"abcdefghij", 4)
    # engine test cut (tests/engine_test.cpp:28,95-96): 43-byte partitions
    data = (b'{"name":"Ada","scores":[1,2],"lang":"en"}\n'
            b'{"name":"Bob","scores":[3,4],"lang":"fr"}\n')
    assert orc.partitions(data, 43) == [(0, 41), (42, 84)]


# ---------------------------------------------------------------------------
# closed forms of the reference tests, restated independently in Python
# ---------------------------------------------------------------------------
def is_escaped(data: bytes, i: int) -> bool:
    # is_escaped_quote, tests/index/string_index_test.cpp:28-34
    k = 0
    while i - k - 1 >= 0 and data[i - k - 1] == 0x5C:
        k += 1
    return k % 2 == 1


def expected_string_bits(data: bytes):
    # tests/index/string_index_test.cpp:36-59
    out, in_string = [], False
    for i, c in enumerate(data):
        if c == 0x22 and not is_escaped(data, i):
            out.append(0 if in_string else 1)
            in_string = not in_string
        else:
            out.append(1 if in_string else 0)
    return out


def expected_leveled_bits(data: bytes, levels: int):
    # tests/index/leveled_bitmap_index_test.cpp:57-92, with the stack depth of
    # the reference's signed char (wraps modulo 256)
    s = expected_string_bits(data)
    rows = [[0] * len(data) for _ in range(levels)]
    level = -1
    for i, c in enumerate(data):
        if s[i]:
            continue
        if c in b"{[":
            level = (level + 1 + 128) % 256 - 128
            if 0 <= level < levels:
                rows[level][i] = 1
        elif c in b"}]":
            if 0 <= level < levels:
                rows[level][i] = 1
            level = (level - 1 + 128) % 256 - 128
        elif c in b":,":
            if 0 <= level < levels:
                rows[level][i] = 1
    return rows


def test_escape_brute_force_all_short_strings(orc):
    # every string of length <= 9 over {\, ", a}
    for n in range(1, 10):
        for tup in itertools.product(b'\\"a', repeat=n):
            data = bytes(tup)
            got = bits(orc.escape_index(data), n)
            want = [1 if is_escaped(data, i) else 0 for i in range(n)]
            assert got == want, data


@pytest.mark.parametrize("seed", range(6))
def test_string_and_leveled_closed_forms(orc, seed):
    data = jsongen.byte_soup(seed, 3000 + 977 * seed)
    assert bits(orc.string_index(data), len(data)) == expected_string_bits(data)
    lv = orc.leveled_index(data, 5)
    want = expected_leveled_bits(data, 5)
    for level in range(5):
        assert bits(lv[level], len(data)) == want[level], level


def test_string_mask_tail_bits_carry_final_state(orc):
    # bits past the end take the in-string state (prefix XOR over zero quotes)
    w = orc.string_index(b'"abc')
    assert int(w[0]) == (1 << 64) - 1
    w = orc.string_index(b'"abc"')
    assert int(w[0]) == 0b01111


def test_leveled_int8_wrap(orc):
    data = b"[" * 258 + b"]" * 258
    lv = orc.leveled_index(data, 3)
    want = expected_leveled_bits(data, 3)
    for level in range(3):
        assert bits(lv[level], len(data)) == want[level]
    # depth 256 wraps to level 0, depth 257 to level 1
    assert bits(lv[0], len(data))[256] == 1 and bits(lv[1], len(data))[257] == 1


def test_newline_index_matches_find(orc):
    data = jsongen.byte_soup(9, 50000)
    assert orc.newline_index(data).tolist() == [0] + [i for i, c in enumerate(data) if c == 10]


def test_chunked_escape_quirk_kat(orc):
    # Hand-derived from escape-carry-index.cu:17-27 / escape-index.cu:18-36:
    # 4 threads over 256 bytes -> 64-byte chunks.  Byte 0 'a', bytes 1..128
    # backslashes, byte 129 '"'.  The carry into chunk 2 is recomputed from 0
    # over the all-backslash chunk 1, so it is 0 instead of the true 1: the
    # reference marks byte 129 escaped and byte 128 not, true semantics the
    # reverse.
    data = b"a" + b"\\" * 128 + b'"' + b"b" * 126
    true = bits(orc.escape_index(data), 256)
    ref = bits(orc.escape_index(data, chunk_threads=4), 256)
    assert true[128] == 1 and true[129] == 0
    assert ref[128] == 0 and ref[129] == 1
    assert true[:128] == ref[:128] and true[130:] == ref[130:]
    # with no 64-byte run of backslashes the two semantics agree
    data = jsongen.ndjson(3, 100_000)
    for threads in (1, 7, 64, 16384):
        assert np.array_equal(orc.escape_index(data), orc.escape_index(data, chunk_threads=threads))


def test_multithreaded_oracle_equals_serial(orc):
    for seed, n in ((1, 100_000), (2, 300_001), (3, 5000)):
        d = jsongen.byte_soup(seed, n)
        a, b = orc.build_index(d, 5, 1), orc.build_index(d, 5, 7)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


# ---------------------------------------------------------------------------
# query interpreter vs an independent DOM evaluator
# ---------------------------------------------------------------------------
class Node:
    def __init__(self, kind, start, end, children=None):
        self.kind, self.start, self.end = kind, start, end
        self.children = children  # object: [(raw_key, Node)], array: [Node]


def parse(data: bytes, i: int = 0):
    """Compact JSON parser recording byte spans of every value."""
    c = data[i]
    if c == 0x7B:
        j, kids = i + 1, []
        if data[j] == 0x7D:
            return Node("obj", i, j + 1, kids), j + 1
        while True:
            assert data[j] == 0x22
            k = j + 1
            while data[k] != 0x22:
                k += 2 if data[k] == 0x5C else 1
            key = data[j + 1:k]
            assert data[k + 1] == 0x3A
            v, j = parse(data, k + 2)
            kids.append((key, v))
            if data[j] == 0x7D:
                return Node("obj", i, j + 1, kids), j + 1
            j += 1
    if c == 0x5B:
        j, kids = i + 1, []
        if data[j] == 0x5D:
            return Node("arr", i, j + 1, kids), j + 1
        while True:
            v, j = parse(data, j)
            kids.append(v)
            if data[j] == 0x5D:
                return Node("arr", i, j + 1, kids), j + 1
            j += 1
    if c == 0x22:
        k = i + 1
        while data[k] != 0x22:
            k += 2 if data[k] == 0x5C else 1
        return Node("str", i, k + 1), k + 1
    k = i
    while k < len(data) and data[k] not in b",]}\n":
        k += 1
    return Node("lit", i, k), k


def dom_eval(data: bytes, node: Node, steps, out):
    """Evaluate a path; returns False when the record is abandoned, which (like
    the reference's goto next_line) drops every later result slot."""
    if not steps:
        out.append((node.start, node.end))
        return True
    kind, arg = steps[0]
    if kind == "key":
        if node.kind != "obj":
            return False
        for key, child in node.children:
            if key == arg:
                return dom_eval(data, child, steps[1:], out)
        return False
    if kind == "idx":
        if node.kind != "arr" or arg >= len(node.children):
            return False
        return dom_eval(data, node.children[arg], steps[1:], out)
    if kind == "ridx":
        if node.kind != "arr" or arg > len(node.children):
            return False
        return dom_eval(data, node.children[len(node.children) - arg], steps[1:], out)
    if kind == "multi":
        for sub in arg:
            if not dom_eval(data, node, [sub] + steps[1:], out):
                return False
        return True
    if kind == "eq":  # [?(@ == "s")]
        if data[node.start:node.end] != b'"' + arg + b'"':
            return False
        return dom_eval(data, node, steps[1:], out)
    if kind == "keyeq":  # [?(@.k == "s")]
        key, val = arg
        if node.kind != "obj":
            return False
        for k, child in node.children:
            if k == key:
                if data[child.start:child.end] != b'"' + val + b'"':
                    return False
                return dom_eval(data, node, steps[1:], out)
        return False
    raise AssertionError(kind)


DOM_QUERIES = [
    ("$.a", [("key", b"a")]),
    ("$.a.b", [("key", b"a"), ("key", b"b")]),
    ("$.tags[1]", [("key", b"tags"), ("idx", 1)]),
    ("$.tags[-1]", [("key", b"tags"), ("ridx", 1)]),
    ("$.tags[-2]", [("key", b"tags"), ("ridx", 2)]),
    ("$.arr[0:3]", [("key", b"arr"), ("multi", [("idx", 0), ("idx", 1), ("idx", 2)])]),
    ("$.arr[2,0]", [("key", b"arr"), ("multi", [("idx", 2), ("idx", 0)])]),
    ("$.x[0].y", [("key", b"x"), ("idx", 0), ("key", b"y")]),
    ('$.k[?(@ == "x")]', [("key", b"k"), ("eq", b"x")]),
    ('$.h[?(@.k == "v")]', [("key", b"h"), ("keyeq", (b"k", b"v"))]),
    ("$['ab']['c']", [("key", b"ab"), ("key", b"c")]),
]


def dom_record(r: random.Random, depth: int) -> str:
    keys = ["a", "b", "c", "k", "x", "y", "h", "ab", "abc", "tags", "arr"]
    roll = r.random()
    if depth >= 4 or roll < 0.35:
        return r.choice(['"x"', '"v"', '"a\\"b"', "12", "true", '"q,:{"', '"w"'])
    if roll < 0.7:
        ks = r.sample(keys, r.randint(1, 4))
        return "{" + ",".join(f'"{k}":{dom_record(r, depth + 1)}' for k in ks) + "}"
    return "[" + ",".join(dom_record(r, depth + 1) for _ in range(r.randint(1, 4))) + "]"


def test_query_oracle_matches_dom_evaluator(orc, gpj):
    r = random.Random(99)
    lines = []
    for _ in range(1500):
        ks = r.sample(["a", "b", "tags", "arr", "x", "k", "h", "ab"], r.randint(1, 5))
        lines.append("{" + ",".join(f'"{k}":{dom_record(r, 1)}' for k in ks) + "}")
    data = ("\n".join(lines)).encode()
    nodes = []
    pos = 0
    for ln in lines:
        nodes.append(parse(data, pos)[0])
        pos += len(ln) + 1
    checked = 0
    for text, steps in DOM_QUERIES:
        ir, depth, slots = gpj.compile_query(text)
        res = orc.query(data, orc.build_index(data, max(depth, 1)), ir, slots)
        for li, node in enumerate(nodes):
            out = []
            dom_eval(data, node, steps, out)
            got = [(int(a), int(b)) for a, b in res[li] if a >= 0]
            assert got == out, (text, lines[li], got, out)
            checked += len(out)
    assert checked > 1000


def test_query_oracle_quirks(orc, gpj):
    # MOVE_TO_KEY falls through on an empty object (query.cu:259-298 with
    # asserts compiled out): $.a.b over {"a":{}} yields an empty span at '}'.
    data = b'{"a":{}}'
    ir, depth, slots = gpj.compile_query("$.a.b")
    res = orc.query(data, orc.build_index(data, 2), ir, slots)
    assert res[0, 0].tolist() == [7, 7]
    # "$" alone stores the root without its closing brace (query.cu:217-229)
    ir, depth, slots = gpj.compile_query("$")
    res = orc.query(data, orc.build_index(data, 1), ir, slots)
    assert data[res[0, 0, 0]:res[0, 0, 1]] == b'{"a":{}'


# ---------------------------------------------------------------------------
# extension semantics ([*], array-element filter) vs the DOM evaluator
# ---------------------------------------------------------------------------
def dom_eval_ext(data: bytes, node, steps, out):
    """Standard JSONPath over arrays: [*] visits every element in order; a
    filter on an array keeps the matching elements, on any other value it
    tests the value itself (the reference's behaviour).  Results in document
    order."""
    if not steps:
        out.append((node.start, node.end))
        return
    kind, arg = steps[0]
    if kind == "each":
        if node.kind == "arr":
            for child in node.children:
                dom_eval_ext(data, child, steps[1:], out)
        return
    if kind == "keyeq":
        key, val = arg
        targets = node.children if node.kind == "arr" else [node]
        for t in targets:
            if t.kind != "obj":
                continue
            for k, child in t.children:
                if k == key:
                    if data[child.start:child.end] == b'"' + val + b'"':
                        dom_eval_ext(data, t, steps[1:], out)
                    break
        return
    if kind == "key":
        if node.kind == "obj":
            for k, child in node.children:
                if k == arg:
                    dom_eval_ext(data, child, steps[1:], out)
                    return
        return
    raise AssertionError(kind)


EXT_DOM_QUERIES = [
    ("$.h[*]", [("key", b"h"), ("each", None)]),
    ("$.h[*].t", [("key", b"h"), ("each", None), ("key", b"t")]),
    ("$.h[*][*]", [("key", b"h"), ("each", None), ("each", None)]),
    ('$.h[?(@.k == "v")].x', [("key", b"h"), ("keyeq", (b"k", b"v")), ("key", b"x")]),
    ('$.h[?(@.k == "v")]', [("key", b"h"), ("keyeq", (b"k", b"v"))]),
    ("$.a.h[*].k", [("key", b"a"), ("key", b"h"), ("each", None), ("key", b"k")]),
]


def test_extension_oracle_matches_dom(orc, gpj):
    r = random.Random(5)
    lines = []
    for _ in range(1200):
        ks = r.sample(["a", "h", "x", "k", "t"], r.randint(1, 4))
        lines.append("{" + ",".join(f'"{k}":{dom_record(r, 1)}' for k in ks) + "}")
    data = ("\n".join(lines)).encode()
    nodes, pos = [], 0
    for ln in lines:
        nodes.append(parse(data, pos)[0])
        pos += len(ln) + 1
    checked = 0
    for text, steps in EXT_DOM_QUERIES:
        ir, depth, _ = gpj.compile_query(text, extensions=True)
        spans, counts = orc.query_ext(data, orc.build_index(data, max(depth, 1)), ir, 64)
        for li, node in enumerate(nodes):
            out = []
            dom_eval_ext(data, node, steps, out)
            got = [(int(a), int(b)) for a, b in spans[li][: int(counts[li])]]
            assert got == out, (text, lines[li], got, out)
            checked += len(out)
    assert checked > 500


def test_extension_mode_keeps_reference_results(orc, gpj):
    # bytecode without extension opcodes gives the same spans in both modes
    data = jsongen.ndjson(8, 200_000, max_depth=5)
    for q in ("$.a", "$.a.b", "$.tags[1]", "$.arr[0:2]"):
        ir, depth, slots = gpj.compile_query(q)
        idx = orc.build_index(data, max(depth, 1))
        fixed = orc.query(data, idx, ir, slots)
        spans, counts = orc.query_ext(data, idx, ir, slots)
        for li in range(fixed.shape[0]):
            got = [tuple(x) for x in spans[li][: int(counts[li])].tolist()]
            want = [tuple(x) for x in fixed[li].tolist() if x[0] >= 0]
            assert got == want


ENGINE_KAT = (b'{"name":"Ada","scores":[1,2],"lang":"en"}\n'
              b'{"name":"Bob","scores":[3,4],"lang":"fr"}\n')


@pytest.mark.parametrize("part_size", [0, 43])
def test_engine_query_reference_kat(orc, gpj, part_size):
    # tests/engine_test.cpp:41-58 (both files of the reference's engine tests;
    # its options cut partitions of 43 bytes, engine_test.cpp:22-28)
    compiled = [gpj.compile_query(q) for q in ("$.name", "$.scores[1]")]
    out = orc.engine_query(ENGINE_KAT, compiled, part_size)
    assert out == [[[b'"Ada"'], [b'"Bob"']], [[b"2"], [b"4"]]]


@pytest.mark.parametrize("seed", [3, 4])
def test_engine_query_partitioning_keeps_line_results(orc, gpj, seed):
    # lines never straddle a partition cut, so the materialized result of the
    # whole file equals that of any partitioning (the cut newline dropped)
    data = jsongen.ndjson(seed, 60000, max_depth=5)
    compiled = [gpj.compile_query(q) for q in ("$.a", "$.user.lang", "$.e[1]", "$.items[0].sku")]
    whole = orc.engine_query(data, compiled, 0)
    assert any(whole)
    for part in (4096, 10007, 33333):
        assert orc.engine_query(data, compiled, part) == whole
    # and each value text is the span the per-line query returned
    ix = orc.build_index(data, 2)
    spans = orc.query(data, ix, compiled[0][0], compiled[0][2])
    texts = [[data[a:b]] for a, b in spans[:, 0] if a >= 0]
    assert whole[0] == texts


def test_key_search_ends_at_a_colon_level_end(orc, gpj):
    # malformed input whose level end is a ':' (a ']' closes the inner object,
    # then a top-level ':'): the key search reaches the level end without a
    # '}' separator; the reference's search_key loop (query.cu:254-302) would
    # test that ':' forever, here the path fails
    data = b'{"a":{"x":1]:2}\n{"a":{"y":5},"b":1}\n'
    ir, depth, slots = gpj.compile_query("$.a.y")
    ix = orc.build_index(data, max(depth, 1))
    got = orc.query(data, ix, ir, slots)
    assert got[0, 0, 0] == -1 and got[0, 0, 1] == -1
    assert data[got[1, 0, 0]:got[1, 0, 1]] == b"5"
