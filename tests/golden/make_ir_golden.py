"""Writes tests/golden/ir_golden.json: bytecode, max depth and slot count (or
the error class) of a list of JSONPath queries, as produced by the reference's
own compiler (src/gpjson/query/query_compiler.cpp, built by `make -C oracle ref`
into oracle/_ref).  Run in a container that has the reference sources; the
fixture is what the tests read."""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle  # noqa: E402

QUERIES = [
    "$", "$.name", "$.a.b", "$.user.lang", "$.e[0].f", "$.store.book[1].author",
    "$.store.magazine", "$.expensive", "$.scores[1]", "$.arr[0:2]", "$.arr[:3]", "$.arr[2:5]",
    "$.arr[-1]", "$.arr[-3]", "$.arr[-2:]", "$.arr[1,3]", "$.arr[0,2,4].x", "$['a']", "$[\"a\"].b",
    "$['a b']['c.d']", "$.a[0][1]", "$.a[1:3].b[0,1]", "$.a[0,1].b[2,3]", "$.a[0:2].b[0:2].c",
    "$.user.lang[?(@ == \"en\")]", "$.user.lang[?(@ == 'nl')]", "$.h[?(@.k == \"v\")].id",
    "$.a[?(@.b.c == \"x\")]", "$.a[?(@ == \"\")]", "$.a[?(@.b==\"x\")].c", "$.a.b ", "$.a[0] ",
    "$.k\\x", "$.a[2:1]", "$.a[0:0]", "$.x[10000]", "$.x[128]", "$.x[300:302]",
    "$.a['q\\'uote']", "$.a[\"dq\\\"x\"]", "$.aaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaaa",
    # malformed
    "$.", "$.arr[-0]", "", "x", "$.a[", "$.a[1", "$.a[x]", "$.a[1,x]", "$.a[:x]", "$.a[-x]",
    "$.a['']", "$.a[1:2", "$[?(@.a != \"x\")]", "$.a[?(@ == x)]", "$.a[12a]", "$.a[?(",
    # unsupported
    "$..name", "$.arr[*]", "$.a[--1]", "$a", "$.a[?(@.b < \"x\")]", "$.a[?(@.b = \"x\")]",
]


def main():
    if not oracle.ref_compiler_available():
        raise SystemExit("oracle/_ref is not built (make -C oracle ref)")
    out = []
    for q in QUERIES:
        try:
            ir, depth, slots = oracle.ref_compile(q)
            out.append({"query": q, "ir": ir.hex(), "max_depth": depth, "slots": slots})
        except ValueError as e:
            kind = str(e).split(":", 1)[0]
            out.append({"query": q, "error": kind, "message": str(e).split(": ", 1)[1]})
    with open(os.path.join(HERE, "ir_golden.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(f"wrote {len(out)} entries")


if __name__ == "__main__":
    main()
