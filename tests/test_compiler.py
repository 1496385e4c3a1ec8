"""The product's JSONPath compiler against the reference compiler's own output
(tests/golden/ir_golden.json, written by tests/golden/make_ir_golden.py from
the reference sources) and against the reference's compiler tests
(tests/query/query_compiler_test.cpp:58-147)."""
import json
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "ir_golden.json")


def golden():
    with open(GOLDEN) as f:
        return json.load(f)


@pytest.mark.parametrize("entry", golden(), ids=lambda e: repr(e["query"])[:40])
def test_matches_reference_compiler(gpj, entry):
    q = entry["query"]
    if "ir" in entry:
        ir, depth, slots = gpj.compile_query(q)
        assert ir.hex() == entry["ir"]
        assert depth == entry["max_depth"]
        assert slots == entry["slots"]
    else:
        cls = {"malformed": gpj.MalformedQueryError,
               "unsupported": gpj.UnsupportedQueryError}[entry["error"]]
        with pytest.raises(cls) as ei:
            gpj.compile_query(q)
        assert str(ei.value) == entry["message"]


def decode(ir: bytes):
    """IRByteInputBuffer-style decoding (src/gpjson/query/ir_input_buffer.cpp:18-45)."""
    out, pos = [], 0

    def varint():
        nonlocal pos
        v, shift = 0, 0
        while True:
            b = ir[pos]
            pos += 1
            v |= (b & 0x7F) << shift
            if not b & 0x80:
                return v
            shift += 7

    while pos < len(ir):
        op = ir[pos]
        pos += 1
        if op in (4, 9):
            n = varint()
            out.append((op, ir[pos:pos + n].decode()))
            pos += n
        elif op in (5, 6):
            out.append((op, varint()))
        else:
            out.append((op, None))
        if op == 0:
            break
    return out


def test_reference_compiler_kats(gpj):
    ir, depth, slots = gpj.compile_query("$.name")
    assert (depth, slots) == (1, 1)
    assert decode(ir) == [(4, "name"), (3, None), (1, None), (0, None)]
    ir, depth, slots = gpj.compile_query("$.arr[0:2]")
    assert (depth, slots) == (2, 2)
    assert decode(ir) == [(4, "arr"), (3, None), (5, 0), (3, None), (1, None), (2, None),
                          (5, 1), (3, None), (1, None), (0, None)]
    ir, depth, slots = gpj.compile_query('$.user.lang[?(@ == "en")]')
    assert (depth, slots) == (2, 1)
    assert decode(ir) == [(4, "user"), (3, None), (4, "lang"), (3, None), (7, None),
                          (9, '"en"'), (8, None), (1, None), (0, None)]


def test_reference_compiler_rejections(gpj):
    for q in ("$.", "$.arr[-0]"):
        with pytest.raises(gpj.MalformedQueryError):
            gpj.compile_query(q)
    for q in ("$..name", "$.arr[*]"):
        with pytest.raises(gpj.UnsupportedQueryError):
            gpj.compile_query(q)


def test_integer_literal_errors(gpj):
    # std::stoull behaviour of the reference (query_compiler.cpp:462-464)
    with pytest.raises(ValueError):
        gpj.compile_query("$.a[1:]")
    with pytest.raises(OverflowError):
        gpj.compile_query("$.a[99999999999999999999999]")


def test_live_reference_compiler_when_built(gpj, orc):
    if not orc.ref_compiler_available():
        pytest.skip("oracle/_ref not built (reference sources absent)")
    import random
    r = random.Random(7)
    atoms = [".a", ".bb", "['c']", "[0]", "[2]", "[-1]", "[1:3]", "[:2]", "[0,2]", "[-2:]",
             "[?(@ == \"x\")]", "[?(@.k == \"v\")]", ".user", "[3,1]"]
    for _ in range(400):
        q = "$" + "".join(r.choice(atoms) for _ in range(r.randint(0, 5)))
        try:
            want = orc.ref_compile(q)
        except ValueError as e:
            kind = str(e).split(":", 1)[0]
            cls = gpj.MalformedQueryError if kind == "malformed" else gpj.UnsupportedQueryError
            with pytest.raises(cls):
                gpj.compile_query(q)
            continue
        assert gpj.compile_query(q) == want, q
