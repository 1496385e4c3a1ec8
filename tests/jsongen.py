"""Seeded random inputs for the tests: NDJSON records with tricky strings,
escape runs and nesting, and raw byte soup dense in the structural bytes."""
from __future__ import annotations

import random

KEYS = ["a", "b", "c", "k", "id", "t", "name", "items", "sku", "user", "lang", "e", "f", "h",
        "x", "y", "tags", "note", "empty", "arr", "nested", "ab", "abc"]


def rand_string(r: random.Random, escapes: bool = True, maxlen: int = 24) -> str:
    n = r.randint(0, maxlen)
    out = []
    pool = 'abcxyz {}[]:,0123456789-_./'
    for _ in range(n):
        if escapes and r.random() < 0.08:
            out.append(r.choice(['\\"', '\\\\', '\\n', '\\t', '\\u00e9', '\\/']))
        else:
            out.append(r.choice(pool))
    return '"' + "".join(out) + '"'


def rand_value(r: random.Random, depth: int, max_depth: int, spaces: bool) -> str:
    sp = (lambda: " " * r.randint(0, 2)) if spaces else (lambda: "")
    roll = r.random()
    if depth >= max_depth or roll < 0.45:
        k = r.random()
        if k < 0.5:
            return rand_string(r)
        if k < 0.8:
            return str(r.randint(-1000, 100000))
        if k < 0.9:
            return f"{r.uniform(-1e3, 1e3):.3f}"
        return r.choice(["true", "false", "null"])
    if roll < 0.75:
        n = r.randint(0, 5)
        keys = r.sample(KEYS, n)
        items = [f"{sp()}\"{k}\"{sp()}:{sp()}{rand_value(r, depth + 1, max_depth, spaces)}{sp()}"
                 for k in keys]
        return "{" + ",".join(items) + "}"
    n = r.randint(0, 5)
    items = [f"{sp()}{rand_value(r, depth + 1, max_depth, spaces)}{sp()}" for _ in range(n)]
    return "[" + ",".join(items) + "]"


def rand_record(r: random.Random, max_depth: int = 5, spaces: bool = False) -> str:
    n = r.randint(0, 7)
    keys = r.sample(KEYS, n)
    sp = (lambda: " " * r.randint(0, 2)) if spaces else (lambda: "")
    items = [f"{sp()}\"{k}\"{sp()}:{sp()}{rand_value(r, 1, max_depth, spaces)}{sp()}" for k in keys]
    return "{" + ",".join(items) + "}"


def ndjson(seed: int, target_bytes: int, max_depth: int = 5, spaces: bool = False) -> bytes:
    r = random.Random(seed)
    parts, size = [], 0
    while size < target_bytes:
        rec = rand_record(r, max_depth, spaces)
        parts.append(rec)
        size += len(rec) + 1
    return ("\n".join(parts) + "\n").encode("latin-1")


def byte_soup(seed: int, n: int, special_weight: float = 0.5) -> bytes:
    """Arbitrary bytes, dense in \\ " { } [ ] : , and newline, plus bytes >= 0x80
    whose low 7 bits alias the special characters."""
    r = random.Random(seed)
    specials = b'\\"{}[]:,\n'
    aliases = bytes(c | 0x80 for c in specials) + bytes([0x20 | 0x80, 0xFF, 0x00])
    out = bytearray()
    while len(out) < n:
        x = r.random()
        if x < special_weight:
            c = r.choice(specials)
            run = r.choice([1, 1, 1, 2, 3, 5, 64, 130]) if c == 0x5C else 1
            out += bytes([c]) * run
        elif x < special_weight + 0.1:
            out.append(r.choice(aliases))
        else:
            out.append(r.randint(0x20, 0x7E))
    return bytes(out[:n])
