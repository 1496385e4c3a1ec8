"""ctypes wrapper over the CPU oracle (oracle/gpj_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the cpu_baseline leg of bench.py, always as the checker or as the timed CPU
baseline, never by the product.

Also holds the engine-level composition of the reference
(src/gpjson/engine.cpp:131-171): partition, index each partition, query,
materialize with global offsets
(src/gpjson/query/kernels/orig/query_executor_host.cu:80-119,
src/gpjson/query/query.cpp:87-109).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgpj_oracle.so")
REF_COMPILER_PATH = os.path.join(HERE, "_ref", "libgpj_refcompiler.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u8p = ctypes.POINTER(ctypes.c_uint8)
        u64p = ctypes.POINTER(ctypes.c_uint64)
        i64p = ctypes.POINTER(ctypes.c_int64)
        u64, i64, cint = ctypes.c_uint64, ctypes.c_int64, ctypes.c_int
        L.gpjo_escape_index.argtypes = [u8p, u64, u64p]
        L.gpjo_escape_index_chunked.argtypes = [u8p, u64, u64, u64p]
        L.gpjo_quote_index.argtypes = [u8p, u64, u64p, u64p]
        L.gpjo_string_index.argtypes = [u64p, u64, u64p]
        L.gpjo_newline_count.argtypes = [u8p, u64]
        L.gpjo_newline_count.restype = u64
        L.gpjo_newline_index.argtypes = [u8p, u64, i64p]
        L.gpjo_newline_index.restype = u64
        L.gpjo_leveled_index.argtypes = [u8p, u64, u64p, cint, u64p]
        L.gpjo_build_index.argtypes = [u8p, u64, cint, u64p, u64p, i64p]
        L.gpjo_build_index_mt.argtypes = [u8p, u64, cint, cint, u64p, u64p, i64p]
        L.gpjo_build_index_mt.restype = u64
        L.gpjo_query.argtypes = [u8p, i64, i64p, i64, u64p, u64p, i64, cint, u8p, i64, cint, i64p]
        L.gpjo_query_mt.argtypes = [u8p, i64, i64p, i64, u64p, u64p, i64, cint, u8p, i64, cint, i64p,
                                    cint]
        L.gpjo_query_ext.argtypes = [u8p, i64, i64p, i64, u64p, u64p, i64, cint, u8p, i64, cint,
                                     i64p, i64p]
        L.gpjo_query_ext_mt.argtypes = [u8p, i64, i64p, i64, u64p, u64p, i64, cint, u8p, i64, cint,
                                        i64p, i64p, cint]
        L.gpjo_partition.argtypes = [u8p, u64, u64, u64p, i64]
        L.gpjo_row_bounds.argtypes = [u64p, i64, i64p, i64]
        L.gpjo_row_bounds.restype = i64
        L.gpjo_query_records.argtypes = [u8p, i64, i64p, i64, u64p, u64p, i64, cint, u8p, i64,
                                         cint, cint, i64p, i64p]
        L.gpjo_partition.restype = i64
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(ctypes.POINTER(t))


def as_bytes(data) -> np.ndarray:
    if isinstance(data, np.ndarray):
        return np.ascontiguousarray(data, dtype=np.uint8)
    if isinstance(data, str):
        data = data.encode("latin-1")
    return np.frombuffer(bytes(data), dtype=np.uint8).copy()


def words(n: int) -> int:
    return (n + 63) // 64


def _buf(arr):
    # ctypes needs a valid pointer even for empty arrays
    return arr if arr.size else np.zeros(1, dtype=arr.dtype)


def escape_index(data, chunk_threads: int | None = None) -> np.ndarray:
    f = as_bytes(data)
    out = np.zeros(max(words(f.size), 1), dtype=np.uint64)
    if chunk_threads is None:
        lib().gpjo_escape_index(_p(_buf(f), ctypes.c_uint8), f.size, _p(out, ctypes.c_uint64))
    else:
        lib().gpjo_escape_index_chunked(_p(_buf(f), ctypes.c_uint8), f.size, chunk_threads,
                                        _p(out, ctypes.c_uint64))
    return out[: words(f.size)]


def quote_index(data, escape=None) -> np.ndarray:
    f = as_bytes(data)
    esc = escape_index(f) if escape is None else escape
    out = np.zeros(max(words(f.size), 1), dtype=np.uint64)
    lib().gpjo_quote_index(_p(_buf(f), ctypes.c_uint8), f.size, _p(_buf(esc), ctypes.c_uint64),
                           _p(out, ctypes.c_uint64))
    return out[: words(f.size)]


def string_index(data) -> np.ndarray:
    q = quote_index(data)
    out = np.zeros(max(q.size, 1), dtype=np.uint64)
    lib().gpjo_string_index(_p(_buf(q), ctypes.c_uint64), q.size, _p(out, ctypes.c_uint64))
    return out[: q.size]


def newline_index(data) -> np.ndarray:
    f = as_bytes(data)
    cnt = lib().gpjo_newline_count(_p(_buf(f), ctypes.c_uint8), f.size)
    out = np.zeros(cnt, dtype=np.int64)
    lib().gpjo_newline_index(_p(_buf(f), ctypes.c_uint8), f.size, _p(out, ctypes.c_int64))
    return out


def leveled_index(data, levels: int, string=None) -> np.ndarray:
    """Level-major int64 words, shape (levels, words)."""
    f = as_bytes(data)
    s = string_index(f) if string is None else string
    w = words(f.size)
    out = np.zeros(max(w * levels, 1), dtype=np.uint64)
    lib().gpjo_leveled_index(_p(_buf(f), ctypes.c_uint8), f.size, _p(_buf(s), ctypes.c_uint64),
                             levels, _p(out, ctypes.c_uint64))
    return out[: w * levels].reshape(levels, w)


def build_index(data, levels: int, threads: int = 1):
    """(string[W], leveled[L,W], newline[1+#nl]) for one partition."""
    f = as_bytes(data)
    w = words(f.size)
    string = np.zeros(max(w, 1), dtype=np.uint64)
    leveled = np.zeros(max(w * levels, 1), dtype=np.uint64)
    cnt = lib().gpjo_newline_count(_p(_buf(f), ctypes.c_uint8), f.size)
    nl = np.zeros(cnt, dtype=np.int64)
    if threads <= 1:
        lib().gpjo_build_index(_p(_buf(f), ctypes.c_uint8), f.size, levels,
                               _p(string, ctypes.c_uint64), _p(leveled, ctypes.c_uint64),
                               _p(nl, ctypes.c_int64))
    else:
        got = lib().gpjo_build_index_mt(_p(_buf(f), ctypes.c_uint8), f.size, levels, threads,
                                        _p(string, ctypes.c_uint64), _p(leveled, ctypes.c_uint64),
                                        _p(nl, ctypes.c_int64))
        assert got == cnt
    return string[:w], leveled[: w * levels].reshape(levels, w), nl


def query(data, index, ir: bytes, num_results: int, threads: int = 1) -> np.ndarray:
    """Result spans int64 [lines, slots, 2] (-1 = no match), partition-local."""
    f = as_bytes(data)
    string, leveled, nl = index
    w = words(f.size)
    res = np.full(max(nl.size * num_results * 2, 1), -1, dtype=np.int64)
    irb = np.frombuffer(bytes(ir), dtype=np.uint8).copy()
    args = (_p(_buf(f), ctypes.c_uint8), f.size, _p(_buf(nl), ctypes.c_int64), nl.size,
            _p(_buf(np.ascontiguousarray(string)), ctypes.c_uint64),
            _p(_buf(np.ascontiguousarray(leveled).reshape(-1)), ctypes.c_uint64), w,
            int(np.asarray(leveled).shape[0]) if np.asarray(leveled).ndim == 2 else 0,
            _p(_buf(irb), ctypes.c_uint8), irb.size, num_results, _p(res, ctypes.c_int64))
    if threads <= 1:
        lib().gpjo_query(*args)
    else:
        lib().gpjo_query_mt(*args, threads)
    return res[: nl.size * num_results * 2].reshape(nl.size, num_results, 2)


def query_ext(data, index, ir: bytes, cap: int = 8, threads: int = 1):
    """Extension mode: (spans int64 [lines, cap, 2], counts int64 [lines]); a
    line's results beyond `cap` are counted but not stored."""
    f = as_bytes(data)
    string, leveled, nl = index
    w = words(f.size)
    res = np.full(max(nl.size * cap * 2, 1), -1, dtype=np.int64)
    counts = np.zeros(max(nl.size, 1), dtype=np.int64)
    irb = np.frombuffer(bytes(ir), dtype=np.uint8).copy()
    lv = np.asarray(leveled)
    args = (_p(_buf(f), ctypes.c_uint8), f.size, _p(_buf(nl), ctypes.c_int64), nl.size,
            _p(_buf(np.ascontiguousarray(string)), ctypes.c_uint64),
            _p(_buf(np.ascontiguousarray(lv).reshape(-1)), ctypes.c_uint64), w,
            int(lv.shape[0]) if lv.ndim == 2 else 0,
            _p(_buf(irb), ctypes.c_uint8), irb.size, cap, _p(res, ctypes.c_int64),
            _p(counts, ctypes.c_int64))
    if threads <= 1:
        lib().gpjo_query_ext(*args)
    else:
        lib().gpjo_query_ext_mt(*args, threads)
    return res[: nl.size * cap * 2].reshape(nl.size, cap, 2), counts[: nl.size]


def row_bounds(index) -> np.ndarray:
    """Positions of every structural character of leveled row 0 (the record
    bounds of a top-level array document: '[', each ',' and ']')."""
    _, leveled, _ = index
    row0 = np.ascontiguousarray(np.asarray(leveled)[0])
    dummy = np.zeros(1, dtype=np.int64)
    cnt = lib().gpjo_row_bounds(_p(_buf(row0), ctypes.c_uint64), row0.size,
                                _p(dummy, ctypes.c_int64), 0)
    out = np.zeros(max(cnt, 1), dtype=np.int64)
    lib().gpjo_row_bounds(_p(_buf(row0), ctypes.c_uint64), row0.size, _p(out, ctypes.c_int64), cnt)
    return out[:cnt]


def query_records(data, index, bounds, ir: bytes, cap: int, variable: bool):
    """Top-level array elements as records: record r spans bounds[r] ..
    bounds[r + 1] and interpreter level k reads leveled row k + 1.  Returns
    (spans int64 [records, cap, 2], counts int64 [records] or None)."""
    f = as_bytes(data)
    string, leveled, _ = index
    w = words(f.size)
    b = np.ascontiguousarray(np.asarray(bounds, dtype=np.int64))
    nrec = max(b.size - 1, 0)
    res = np.full(max(nrec * cap * 2, 1), -1, dtype=np.int64)
    counts = np.zeros(max(nrec, 1), dtype=np.int64)
    irb = np.frombuffer(bytes(ir), dtype=np.uint8).copy()
    lv = np.asarray(leveled)
    lib().gpjo_query_records(_p(_buf(f), ctypes.c_uint8), f.size, _p(_buf(b), ctypes.c_int64),
                             b.size, _p(_buf(np.ascontiguousarray(string)), ctypes.c_uint64),
                             _p(_buf(np.ascontiguousarray(lv).reshape(-1)), ctypes.c_uint64), w,
                             int(lv.shape[0]), _p(_buf(irb), ctypes.c_uint8), irb.size, cap,
                             int(bool(variable)), _p(res, ctypes.c_int64),
                             _p(counts, ctypes.c_int64))
    spans = res[: nrec * cap * 2].reshape(nrec, cap, 2)
    return spans, (counts[:nrec] if variable else None)


class PartitionError(Exception):
    pass


def partitions(data, part_size: int):
    """[(start, end)] as create_partitions_for_bytes (file_reader.cpp:41-77)."""
    f = as_bytes(data)
    cap = 16
    while True:
        b = np.zeros(2 * cap, dtype=np.uint64)
        k = lib().gpjo_partition(_p(_buf(f), ctypes.c_uint8), f.size, part_size,
                                 _p(b, ctypes.c_uint64), cap)
        if k < 0:
            raise PartitionError(f"no newline before cut (size {part_size})")
        if k <= cap:
            return [(int(b[2 * i]), int(b[2 * i + 1])) for i in range(k)]
        cap = k


def engine_query(data, compiled, part_size: int = 0, threads: int = 1):
    """Engine::query over an in-memory file: per query, the list of matched
    lines, each a list of value texts (bytes) -- the MaterializedBatchResult.
    `compiled` is a list of (ir_bytes, max_depth, slots)."""
    f = as_bytes(data)
    levels = max([1] + [c[1] for c in compiled])
    out = [[] for _ in compiled]
    for (s, e) in partitions(f, part_size):
        part = f[s:e]
        if part.size == 0:
            continue
        idx = build_index(part, levels, threads)
        for qi, (ir, _depth, slots) in enumerate(compiled):
            res = query(part, idx, ir, slots, threads)
            for line in range(res.shape[0]):
                vals = []
                for slot in range(slots):
                    a, b = int(res[line, slot, 0]), int(res[line, slot, 1])
                    if a < 0 or b < 0:
                        continue
                    vals.append(part[a:b].tobytes())
                if vals:
                    out[qi].append(vals)
    return out


# ---------------------------------------------------------------------------
# Reference compiler (oracle/_ref, built from the reference's own sources)
# ---------------------------------------------------------------------------
_ref = None


def ref_compiler_available() -> bool:
    return os.path.exists(REF_COMPILER_PATH)


def ref_compile(query_text: str):
    """(ir_bytes, max_depth, slots) or raises ValueError('malformed'|'unsupported')."""
    global _ref
    if _ref is None:
        _ref = ctypes.CDLL(REF_COMPILER_PATH)
        _ref.gpj_ref_compile.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint8), ctypes.c_long,
                                         ctypes.POINTER(ctypes.c_long), ctypes.POINTER(ctypes.c_long),
                                         ctypes.c_char_p, ctypes.c_long]
        _ref.gpj_ref_compile.restype = ctypes.c_long
    buf = (ctypes.c_uint8 * 65536)()
    depth, slots = ctypes.c_long(), ctypes.c_long()
    err = ctypes.create_string_buffer(512)
    n = _ref.gpj_ref_compile(query_text.encode("latin-1"), buf, 65536, ctypes.byref(depth),
                             ctypes.byref(slots), err, 512)
    if n == -1:
        raise ValueError("malformed: " + err.value.decode("latin-1"))
    if n == -2:
        raise ValueError("unsupported: " + err.value.decode("latin-1"))
    if n < 0:
        raise ValueError("error: " + err.value.decode("latin-1"))
    return bytes(buf[:n]), depth.value, slots.value
