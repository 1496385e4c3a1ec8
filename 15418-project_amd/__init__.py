"""Python bindings (ctypes) over the C ABI of the MI355X GpJSON engine.

The product is the HIP library `_build/libgpj_hip.so` (include/gpj_capi.h) and
the C++ host API built on it.  This module is thin glue for tests and
benchmarks: it loads the in-tree library, maps status codes onto the
reference's error classes and wraps device buffers.  There is no fallback:
if the HIP library is missing or no GPU is present, device calls raise.

Load it with `load_package()` from __graft_entry__ (the directory name is not
a Python identifier), e.g.
    gpj = __graft_entry__.load_package()
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "_build", "libgpj_hip.so")
REPO_ROOT = os.path.dirname(PKG_DIR)
CAPI_HEADER = os.path.join(REPO_ROOT, "include", "gpj_capi.h")

OK = 0
ERR_INVALID_ARGUMENT = -1
ERR_DEVICE = -2
ERR_OUT_OF_MEMORY = -3
ERR_QUERY_EXECUTION = -4
ERR_MALFORMED_QUERY = -5
ERR_UNSUPPORTED_QUERY = -6
ERR_CAPACITY = -7
ERR_OUT_OF_RANGE = -8
COMPILE_EXTENSIONS = 1
MAX_QUERY_IR_BYTES = 4096
MAX_QUERY_LEVELS = 16
MAX_INDEX_LEVELS = 22
STATIC_MAX_BYTES = 16 << 30  # GPJ_STATIC_MAX_BYTES: inputs of gpj_index_build_static


class GpJSONError(Exception):
    """Base of the reference's error tree (include/gpjson/error/error.hpp:9-17)."""


class ImplementationError(GpJSONError):
    pass


class DeviceError(GpJSONError):
    """CudaError in the reference (src/gpjson/cuda/cuda.cpp:19-29)."""


class MalformedQueryError(GpJSONError):
    pass


class UnsupportedQueryError(GpJSONError):
    pass


class QueryExecutionError(GpJSONError):
    pass


class CapacityError(GpJSONError):
    pass


class IndexT(ctypes.Structure):
    _fields_ = [
        ("d_newline", ctypes.c_void_p),
        ("d_num_lines", ctypes.c_void_p),
        ("num_lines", ctypes.c_int64),
        ("d_string", ctypes.c_void_p),
        ("d_leveled", ctypes.c_void_p),
        ("level_size", ctypes.c_int64),
        ("num_levels", ctypes.c_int32),
        ("owns_memory", ctypes.c_int32),
    ]


class QueryDescT(ctypes.Structure):
    _fields_ = [
        ("d_ir", ctypes.c_void_p),
        ("ir_len", ctypes.c_uint32),
        ("capacity", ctypes.c_int32),
        ("variable", ctypes.c_int32),
        ("max_level", ctypes.c_int32),
        ("d_spans", ctypes.c_void_p),
        ("d_counts", ctypes.c_void_p),
    ]


class DenseResultT(ctypes.Structure):
    _fields_ = [
        ("d_offsets", ctypes.c_void_p),
        ("d_spans", ctypes.c_void_p),
        ("capacity", ctypes.c_int64),
    ]


class CarryT(ctypes.Structure):
    """gpj_carry: the index scan's state before a shard."""
    _fields_ = [("escape", ctypes.c_int32), ("in_string", ctypes.c_int32),
                ("depth", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("newlines", ctypes.c_int64)]


class ShardSummaryT(ctypes.Structure):
    """gpj_shard_summary: a shard's transfer table and newline count."""
    _fields_ = [("table", ctypes.c_uint64), ("newlines", ctypes.c_uint64)]


class BatchFileT(ctypes.Structure):
    _fields_ = [
        ("offset", ctypes.c_uint64),
        ("bytes", ctypes.c_uint64),
        ("newline_offset", ctypes.c_uint64),
        ("newline_capacity", ctypes.c_uint64),
        ("query_mask", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]


class BatchResultsT(ctypes.Structure):
    _fields_ = [
        ("d_spans", ctypes.c_void_p * 2),
        ("d_counts", ctypes.c_void_p * 2),
        ("capacity", ctypes.c_int32),
    ]


BATCH_ALIGN = 16384
BATCH_SLOTS = 2


class IndexOptsT(ctypes.Structure):
    _fields_ = [
        ("grid_size", ctypes.c_int32),
        ("block_size", ctypes.c_int32),
        ("reduction_grid_size", ctypes.c_int32),
        ("reduction_block_size", ctypes.c_int32),
        ("newline_capacity", ctypes.c_int64),
    ]


_lib = None


def lib():
    """The loaded C ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        vp, u64, i64, i32, u32 = (ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64,
                                  ctypes.c_int32, ctypes.c_uint32)
        u32p = ctypes.POINTER(ctypes.c_uint32)
        L.gpj_last_error.restype = ctypes.c_char_p
        L.gpj_device_available.restype = ctypes.c_int
        L.gpj_alloc.argtypes = [ctypes.POINTER(vp), u64]
        L.gpj_free.argtypes = [vp]
        L.gpj_h2d.argtypes = [vp, vp, u64, vp]
        L.gpj_d2h.argtypes = [vp, vp, u64, vp]
        L.gpj_stream_sync.argtypes = [vp]
        L.gpj_index_build.argtypes = [vp, u64, i32, ctypes.POINTER(IndexOptsT),
                                      ctypes.POINTER(IndexT), vp]
        L.gpj_index_free.argtypes = [ctypes.POINTER(IndexT)]
        L.gpj_index_scratch_bytes.argtypes = [u64]
        L.gpj_index_scratch_bytes.restype = u64
        L.gpj_index_build_into.argtypes = [vp, u64, i32, vp, vp, vp, i64, vp, vp, vp]
        L.gpj_index_build_shard.argtypes = [vp, u64, i32, vp, vp, vp, i64, vp, vp,
                                            ctypes.POINTER(CarryT), vp]
        L.gpj_index_shard_summary.argtypes = [vp, u64, vp, vp]
        L.gpj_shard_carries.argtypes = [ctypes.POINTER(ShardSummaryT), i32,
                                        ctypes.POINTER(CarryT)]
        L.gpj_index_status.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_int32), vp]
        L.gpj_index_build_static.argtypes = [vp, u64, i32, vp, vp, vp, i64, vp, vp]
        L.gpj_debug_set_wait_limit.argtypes = [i64]
        L.gpj_debug_escape_quote.argtypes = [vp, u64, vp, vp, vp]
        L.gpj_debug_copy.argtypes = [vp, vp, u64, vp]
        L.gpj_debug_set_narrow_span.argtypes = [i64]
        L.gpj_debug_set_index_pairing.argtypes = [i64]
        L.gpj_query_set_walk.argtypes = [i32]
        L.gpj_query_run.argtypes = [vp, u64, ctypes.POINTER(IndexT), vp, u32, i32, i64, vp, vp]
        L.gpj_query_run_ext.argtypes = [vp, u64, ctypes.POINTER(IndexT), vp, u32, i32, i64, vp,
                                        vp, vp]
        L.gpj_query_run_batch.argtypes = [vp, u64, ctypes.POINTER(IndexT), ctypes.POINTER(QueryDescT),
                                          i32, i64, vp]
        L.gpj_query_compile.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_uint8), u32,
                                        u32p, u32p, u32p, u32]
        L.gpj_query_dense_scratch_bytes.argtypes = [i64, i32, i64]
        L.gpj_query_dense_scratch_bytes.restype = u64
        L.gpj_query_run_dense.argtypes = [vp, u64, ctypes.POINTER(IndexT),
                                          ctypes.POINTER(QueryDescT), i32, i64,
                                          ctypes.POINTER(DenseResultT), vp, vp]
        L.gpj_query_run_records_dense.argtypes = [vp, u64, ctypes.POINTER(IndexT), vp, vp,
                                                  ctypes.POINTER(QueryDescT), i32, i64,
                                                  ctypes.POINTER(DenseResultT), vp, vp]
        L.gpj_row_bounds_scratch_bytes.argtypes = [i64]
        L.gpj_row_bounds_scratch_bytes.restype = u64
        L.gpj_row_bounds.argtypes = [ctypes.POINTER(IndexT), vp, i64, vp, vp, vp]
        L.gpj_query_run_records.argtypes = [vp, u64, ctypes.POINTER(IndexT), vp, vp,
                                            ctypes.POINTER(QueryDescT), i32, i64, vp]
        L.gpj_batch_scratch_bytes.argtypes = [u64]
        L.gpj_batch_scratch_bytes.restype = u64
        L.gpj_batch_tile_map.argtypes = [ctypes.POINTER(BatchFileT), i32, u64, vp]
        L.gpj_batch_index.argtypes = [vp, u64, vp, i32, vp, i32, vp, vp, vp, vp, vp, vp]
        L.gpj_batch_line_counts.argtypes = [vp, u64, vp, i32, vp, vp, vp, vp]
        L.gpj_gather_words.argtypes = [vp, ctypes.POINTER(vp), i32, vp]
        L.gpj_batch_query.argtypes = [vp, u64, vp, i32, i32, vp, vp, vp, vp, i64,
                                      ctypes.POINTER(QueryDescT), i32,
                                      ctypes.POINTER(BatchResultsT), vp]
        L.gpj_batch_query_shaped.argtypes = [vp, u64, vp, i32, i32, vp, vp, vp, vp, i64, i32,
                                             ctypes.POINTER(QueryDescT), i32,
                                             ctypes.POINTER(BatchResultsT), vp]
        L.gpj_batch_per_file.argtypes = [vp, ctypes.POINTER(BatchFileT), i32, i32, vp, vp, vp,
                                         vp, vp, ctypes.POINTER(QueryDescT), i32,
                                         ctypes.POINTER(BatchResultsT), vp, vp]
        L.gpj_stream_create.argtypes = [ctypes.POINTER(vp)]
        L.gpj_stream_destroy.argtypes = [vp]
        L.gpj_graph_begin.argtypes = [vp]
        L.gpj_graph_end.argtypes = [vp, ctypes.POINTER(vp)]
        L.gpj_graph_launch.argtypes = [vp, vp]
        L.gpj_graph_destroy.argtypes = [vp]
        for name in ("gpj_alloc", "gpj_free", "gpj_h2d", "gpj_d2h", "gpj_stream_sync",
                     "gpj_device_sync", "gpj_index_build", "gpj_index_free",
                     "gpj_index_build_into", "gpj_index_status", "gpj_debug_set_wait_limit",
                     "gpj_index_build_static",
                     "gpj_debug_escape_quote", "gpj_query_run",
                     "gpj_query_run_ext", "gpj_query_run_batch", "gpj_query_compile",
                     "gpj_debug_copy", "gpj_row_bounds", "gpj_query_run_records",
                     "gpj_query_run_dense", "gpj_query_run_records_dense",
                     "gpj_batch_tile_map", "gpj_batch_index", "gpj_batch_query",
                     "gpj_batch_line_counts", "gpj_gather_words",
                     "gpj_batch_query_shaped",
                     "gpj_batch_per_file", "gpj_graph_begin", "gpj_graph_end",
                     "gpj_graph_launch", "gpj_graph_destroy", "gpj_stream_create",
                     "gpj_stream_destroy", "gpj_debug_set_narrow_span", "gpj_query_set_walk",
                     "gpj_debug_set_index_pairing",
                     "gpj_index_build_shard", "gpj_index_shard_summary", "gpj_shard_carries"):
            getattr(L, name).restype = i32
        _lib = L
    return _lib


def check(status: int, what: str = "") -> None:
    if status == OK:
        return
    msg = lib().gpj_last_error().decode("utf-8", "replace")
    if what:
        msg = f"{what}: {msg}"
    cls = {
        ERR_MALFORMED_QUERY: MalformedQueryError,
        ERR_UNSUPPORTED_QUERY: UnsupportedQueryError,
        ERR_QUERY_EXECUTION: QueryExecutionError,
        ERR_DEVICE: DeviceError,
        ERR_OUT_OF_MEMORY: DeviceError,
        ERR_CAPACITY: CapacityError,
        ERR_OUT_OF_RANGE: OverflowError,
        ERR_INVALID_ARGUMENT: ValueError,
    }.get(status, GpJSONError)
    raise cls(msg)


def new_stream():
    """A new HIP stream (default priority, non-blocking) as an int handle."""
    s = ctypes.c_void_p()
    check(lib().gpj_stream_create(ctypes.byref(s)), "stream create")
    return s.value


def device_available() -> bool:
    return bool(lib().gpj_device_available())


def exported_symbols_declared_in_header() -> list[str]:
    """Function names declared in include/gpj_capi.h."""
    import re
    text = open(CAPI_HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?[a-z_0-9]+\s*\*?\s*(gpj_[a-z_0-9]+)\s*\(",
                                 text, flags=re.M)))


# ---------------------------------------------------------------------------
# compiler
# ---------------------------------------------------------------------------
def compile_query(query: str, extensions: bool = False):
    """(ir_bytes, max_depth, num_slots); raises MalformedQueryError etc."""
    buf = (ctypes.c_uint8 * 65536)()
    n, d, s = ctypes.c_uint32(), ctypes.c_uint32(), ctypes.c_uint32()
    st = lib().gpj_query_compile(query.encode("latin-1"), buf, 65536, ctypes.byref(n),
                                 ctypes.byref(d), ctypes.byref(s),
                                 COMPILE_EXTENSIONS if extensions else 0)
    check(st)
    return bytes(buf[: n.value]), d.value, s.value


def uses_extension_opcodes(ir: bytes) -> bool:
    """True if the bytecode holds MOVE_TO_EACH (10) or JUMP_IF_NOT_ARRAY (11),
    i.e. it must run through the variable-length entry point."""
    pos = 0

    def varint():
        nonlocal pos
        v, shift = 0, 0
        while pos < len(ir):
            b = ir[pos]
            pos += 1
            v |= (b & 0x7F) << shift
            if not b & 0x80:
                break
            shift += 7
        return v

    while pos < len(ir):
        op = ir[pos]
        pos += 1
        if op in (10, 11):
            return True
        if op in (4, 9):
            n = varint()
            pos += n
        elif op in (5, 6):
            varint()
        elif op == 0:
            continue
    return False


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class DeviceBuffer:
    """Owning device allocation (DeviceArray, include/gpjson/cuda/cuda.hpp:10-40)."""

    def __init__(self, nbytes: int):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(lib().gpj_alloc(ctypes.byref(p), self.nbytes), "alloc")
        self.ptr = p.value or 0

    @classmethod
    def from_host(cls, arr, stream=None) -> "DeviceBuffer":
        a = np.ascontiguousarray(arr)
        b = cls(max(a.nbytes, 1))
        if a.nbytes:
            check(lib().gpj_h2d(b.ptr, a.ctypes.data, a.nbytes, stream), "h2d")
            check(lib().gpj_stream_sync(stream))
        return b

    def to_host(self, dtype, count: int | None = None, stream=None) -> np.ndarray:
        dt = np.dtype(dtype)
        cnt = self.nbytes // dt.itemsize if count is None else int(count)
        out = np.empty(cnt, dtype=dt)
        if cnt:
            check(lib().gpj_d2h(out.ctypes.data, self.ptr, cnt * dt.itemsize, stream), "d2h")
        return out

    def free(self):
        if self.ptr:
            check(lib().gpj_free(self.ptr))
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def d2h(ptr: int, dtype, count: int, stream=None) -> np.ndarray:
    dt = np.dtype(dtype)
    out = np.empty(int(count), dtype=dt)
    if count:
        check(lib().gpj_d2h(out.ctypes.data, ptr, out.nbytes, stream), "d2h")
    return out


# ---------------------------------------------------------------------------
# index + query
# ---------------------------------------------------------------------------
class BuiltIndex:
    """Owning wrapper of gpj_index (BuiltIndices, include/gpjson/index/index.hpp:83-109)."""

    def __init__(self, raw: IndexT, n: int):
        self.raw = raw
        self.n = n

    @property
    def num_lines(self) -> int:
        return int(self.raw.num_lines)

    @property
    def level_size(self) -> int:
        return int(self.raw.level_size)

    @property
    def num_levels(self) -> int:
        return int(self.raw.num_levels)

    def newline(self) -> np.ndarray:
        return d2h(self.raw.d_newline, np.int64, self.num_lines)

    def string(self) -> np.ndarray:
        return d2h(self.raw.d_string, np.uint64, self.level_size)

    def leveled(self) -> np.ndarray:
        w = self.level_size
        return d2h(self.raw.d_leveled, np.uint64, w * self.num_levels).reshape(self.num_levels, w)

    def free(self):
        if self.raw.owns_memory:
            lib().gpj_index_free(ctypes.byref(self.raw))

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def index_status(scratch: DeviceBuffer | int, n: int, stream=None) -> None:
    """Raises DeviceError when the last index build over n bytes that used
    this scratch failed (gpj_index_status: a tile's bounded wait gave up)."""
    ptr = scratch.ptr if isinstance(scratch, DeviceBuffer) else int(scratch)
    failed = ctypes.c_int32()
    check(lib().gpj_index_status(ptr, int(n), ctypes.byref(failed), stream), "index status")
    if failed.value:
        raise DeviceError("index build: a tile's wait for its carry timed out")


class wait_limit:
    """Context manager: index tiles give up after `polls` unsuccessful polls
    for their carry (default 2^24); 0 makes every waiting tile give up at once,
    which drives the failure path without any kernel hanging."""

    def __init__(self, polls: int):
        self.polls = polls

    def __enter__(self):
        check(lib().gpj_debug_set_wait_limit(self.polls), "wait limit")
        return self

    def __exit__(self, *exc):
        check(lib().gpj_debug_set_wait_limit(-1), "wait limit")
        return False


def build_index(dev: DeviceBuffer | int, n: int, levels: int, newline_capacity: int = 0,
                stream=None) -> BuiltIndex:
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    raw = IndexT()
    opts = IndexOptsT(0, 0, 0, 0, int(newline_capacity))
    check(lib().gpj_index_build(ptr, int(n), int(levels), ctypes.byref(opts), ctypes.byref(raw),
                                stream), "index build")
    return BuiltIndex(raw, n)


class ShardIndex:
    """A stream-ordered index of one shard (gpj_index_build_into, or
    gpj_index_build_shard from a carry-in) with its own buffers."""

    def __init__(self, dev: DeviceBuffer | int, n: int, levels: int, carry: CarryT | None = None,
                 newline_capacity: int | None = None, stream=None):
        L = lib()
        ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
        w = (n + 63) // 64
        cap = n + 2 if newline_capacity is None else int(newline_capacity)
        self.n, self.levels, self.words, self.cap = n, levels, w, cap
        self.string = DeviceBuffer(max(w * 8, 8))
        self.leveled = DeviceBuffer(max(w * 8 * levels, 8))
        self.newline = DeviceBuffer(max(cap * 8, 8))
        self.lines = DeviceBuffer(8)
        self.scratch = DeviceBuffer(L.gpj_index_scratch_bytes(n))
        if carry is None:
            check(L.gpj_index_build_into(ptr, n, levels, self.string.ptr, self.leveled.ptr,
                                         self.newline.ptr, cap, self.lines.ptr, self.scratch.ptr,
                                         stream), "shard index")
        else:
            check(L.gpj_index_build_shard(ptr, n, levels, self.string.ptr, self.leveled.ptr,
                                          self.newline.ptr, cap, self.lines.ptr,
                                          self.scratch.ptr, ctypes.byref(carry), stream),
                  "shard index from a carry-in")
        check(L.gpj_stream_sync(stream))

    def summary(self, stream=None) -> ShardSummaryT:
        d = DeviceBuffer(16)
        check(lib().gpj_index_shard_summary(self.scratch.ptr, self.n, d.ptr, stream), "summary")
        check(lib().gpj_stream_sync(stream))
        v = d.to_host(np.uint64, 2, stream)
        return ShardSummaryT(int(v[0]), int(v[1]))

    def string_words(self) -> np.ndarray:
        return self.string.to_host(np.uint64, self.words)

    def leveled_words(self) -> np.ndarray:
        return self.leveled.to_host(np.uint64, self.words * self.levels).reshape(self.levels, -1)

    def newlines(self) -> np.ndarray:
        k = int(self.lines.to_host(np.int64, 1)[0])
        if k < 0:
            raise DeviceError("shard index build timed out")
        return self.newline.to_host(np.int64, min(k, self.cap))


def shard_carries(summaries) -> list:
    """The state before each shard and after the last (gpj_shard_carries)."""
    arr = (ShardSummaryT * max(len(summaries), 1))(*summaries)
    out = (CarryT * (len(summaries) + 1))()
    check(lib().gpj_shard_carries(arr, len(summaries), out), "shard carries")
    return list(out)


def run_query(dev: DeviceBuffer | int, n: int, index: BuiltIndex, ir: bytes, num_results: int,
              stream=None) -> np.ndarray:
    """Spans int64 [num_lines, num_results, 2] (-1 = no match), partition-local."""
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    lines = index.num_lines
    d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
    spans = DeviceBuffer(max(lines * num_results * 16, 16))
    check(lib().gpj_query_run(ptr, int(n), ctypes.byref(index.raw), d_ir.ptr, len(ir),
                              int(num_results), lines, spans.ptr, stream), "query")
    check(lib().gpj_stream_sync(stream))
    out = spans.to_host(np.int64, lines * num_results * 2, stream)
    return out.reshape(lines, num_results, 2)


def run_query_ext(dev: DeviceBuffer | int, n: int, index: BuiltIndex, ir: bytes,
                  capacity: int = 8, stream=None):
    """Extension bytecode: (spans int64 [lines, capacity, 2], counts int64 [lines]).
    Re-runs with the exact capacity when a line has more results."""
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    lines = index.num_lines
    d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
    while True:
        spans = DeviceBuffer(max(lines * capacity * 16, 16))
        counts = DeviceBuffer(max(lines * 8, 8))
        check(lib().gpj_query_run_ext(ptr, int(n), ctypes.byref(index.raw), d_ir.ptr, len(ir),
                                      int(capacity), lines, spans.ptr, counts.ptr, stream),
              "query")
        check(lib().gpj_stream_sync(stream))
        c = counts.to_host(np.int64, lines, stream)
        need = int(c.max()) if c.size else 0
        if need <= capacity:
            sp = spans.to_host(np.int64, lines * capacity * 2, stream).reshape(lines, capacity, 2)
            # entries past a line's count are unspecified: present them as -1
            sp[np.arange(capacity)[None, :] >= c[:, None]] = -1
            return sp, c
        capacity = need


def run_query_batch(dev: DeviceBuffer | int, n: int, index: BuiltIndex, compiled,
                    capacity: int = 8, stream=None):
    """Run several compiled queries in one launch.  `compiled`: list of
    (ir, max_depth, slots).  Returns, per query, spans [lines, cap, 2] (-1 =
    none) and, for extension bytecode, counts (else None)."""
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    lines = index.num_lines
    keep, descs = [], (QueryDescT * max(len(compiled), 1))()
    for i, (ir, depth, slots) in enumerate(compiled):
        var = uses_extension_opcodes(ir)
        cap = capacity if var else slots
        d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
        sp = DeviceBuffer(max(lines * cap * 16, 16))
        ct = DeviceBuffer(max(lines * 8, 8)) if var else None
        keep.append((d_ir, sp, ct, cap, var))
        descs[i] = QueryDescT(d_ir.ptr, len(ir), cap, int(var), int(depth), sp.ptr,
                              ct.ptr if ct else None)
    check(lib().gpj_query_run_batch(ptr, int(n), ctypes.byref(index.raw), descs, len(compiled),
                                    lines, stream), "query batch")
    check(lib().gpj_stream_sync(stream))
    out = []
    for d_ir, sp, ct, cap, var in keep:
        spans = sp.to_host(np.int64, lines * cap * 2, stream).reshape(lines, cap, 2)
        counts = None
        if var:
            counts = ct.to_host(np.int64, lines, stream)
            spans[np.arange(cap)[None, :] >= counts[:, None]] = -1
        out.append((spans, counts))
    return out


def run_query_dense(dev: DeviceBuffer | int, n: int, index: BuiltIndex, compiled,
                    capacity: int | None = None, bounds=None, num_bounds: int = 0,
                    max_records: int | None = None, stream=None, rerun: bool = True):
    """Dense results (gpj_query_run_dense, or gpj_query_run_records_dense
    over top-level array elements when `bounds` is given): per query, CSR
    offsets int64 [records + 1] and spans int64 [total, 2] in record order.
    `capacity` spans per query (default: 2 per record); when a total exceeds
    it, the call reruns with the total (rerun=False returns the truncated
    output and the offsets instead)."""
    L = lib()
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    if bounds is None:
        records = index.num_lines if max_records is None else int(max_records)
    else:
        records = max(num_bounds - 1, 0) if max_records is None else int(max_records)
    nq = len(compiled)
    descs = (QueryDescT * max(nq, 1))()
    irs = []
    for i, (ir, depth, slots) in enumerate(compiled):
        d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
        irs.append(d_ir)
        descs[i] = QueryDescT(d_ir.ptr, len(ir), slots, int(uses_extension_opcodes(ir)),
                              int(depth), None, None)
    caps = [2 * records + 1 if capacity is None else int(capacity)] * nq
    d_m = None
    if bounds is not None:
        d_m = DeviceBuffer.from_host(np.array([num_bounds], dtype=np.int64), stream)
    while True:
        offs = [DeviceBuffer((records + 1) * 8) for _ in range(nq)]
        spans = [DeviceBuffer(max(c * 16, 16)) for c in caps]
        res = (DenseResultT * max(nq, 1))()
        for i in range(nq):
            res[i] = DenseResultT(offs[i].ptr, spans[i].ptr, caps[i])
        scratch = DeviceBuffer(L.gpj_query_dense_scratch_bytes(records, nq, sum(caps)))
        if bounds is None:
            check(L.gpj_query_run_dense(ptr, int(n), ctypes.byref(index.raw), descs, nq, records,
                                        res, scratch.ptr, stream), "dense query")
        else:
            d_b = bounds.ptr if isinstance(bounds, DeviceBuffer) else int(bounds)
            check(L.gpj_query_run_records_dense(ptr, int(n), ctypes.byref(index.raw), d_b,
                                                d_m.ptr, descs, nq, records, res, scratch.ptr,
                                                stream), "dense record query")
        check(L.gpj_stream_sync(stream))
        out = [o.to_host(np.int64, records + 1, stream) for o in offs]
        totals = [int(o[-1]) for o in out]
        if rerun and any(t > c for t, c in zip(totals, caps)):
            caps = [max(t, c) for t, c in zip(totals, caps)]
            continue
        return [(o, sp.to_host(np.int64, 2 * min(t, c), stream).reshape(-1, 2))
                for o, sp, t, c in zip(out, spans, totals, caps)]


def dense_from_slots(spans: np.ndarray, counts: np.ndarray | None = None):
    """CSR form (offsets, spans [total, 2]) of per-record results
    [records][cap][2]: the first counts[r] entries of record r (fixed mode:
    its stored slots, those with start >= 0)."""
    if counts is None:
        counts = (spans[:, :, 0] >= 0).sum(axis=1)
    counts = np.asarray(counts, dtype=np.int64)
    offsets = np.zeros(counts.size + 1, dtype=np.int64)
    np.cumsum(counts, out=offsets[1:])
    keep = np.arange(spans.shape[1])[None, :] < counts[:, None]
    return offsets, spans[keep].reshape(-1, 2)


def compile_element_query(query: str):
    """Top-level-array mode: a query `$[*]<rest>` runs as `$<rest>` on every
    element of the document's top-level array.  Returns the compiled rest
    (ir, max_depth, slots); the index needs max_depth + 1 leveled rows."""
    q = query.strip()
    if not q.startswith("$[*]"):
        raise MalformedQueryError(f"array-element query must start with $[*]: {query!r}")
    return compile_query("$" + q[4:], extensions=True)


def array_bounds(index: BuiltIndex, stream=None):
    """Record bounds of a top-level array document: (DeviceBuffer of int64
    positions, count M); there are M - 1 element records."""
    L = lib()
    w = index.level_size
    scratch = DeviceBuffer(L.gpj_row_bounds_scratch_bytes(w))
    d_count = DeviceBuffer(8)
    # first pass learns M, second writes (capacity known)
    cap = 1
    while True:
        bounds = DeviceBuffer(max(cap * 8, 8))
        check(L.gpj_row_bounds(ctypes.byref(index.raw), bounds.ptr, cap, d_count.ptr,
                               scratch.ptr, stream), "row bounds")
        check(L.gpj_stream_sync(stream))
        m = int(d_count.to_host(np.int64, 1, stream)[0])
        if m <= cap:
            return bounds, m
        cap = m


def run_query_records(dev: DeviceBuffer | int, n: int, index: BuiltIndex, bounds, num_bounds: int,
                      compiled, capacity: int = 8, stream=None):
    """Run compiled element queries (see compile_element_query) over the
    top-level array's elements in one launch.  Returns, per query, spans
    [records, cap, 2] (-1 = none) and counts (None for fixed-slot bytecode)."""
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    records = max(num_bounds - 1, 0)
    keep, descs = [], (QueryDescT * max(len(compiled), 1))()
    for i, (ir, depth, slots) in enumerate(compiled):
        var = uses_extension_opcodes(ir)
        cap = capacity if var else slots
        d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
        sp = DeviceBuffer(max(records * cap * 16, 16))
        ct = DeviceBuffer(max(records * 8, 8)) if var else None
        keep.append((d_ir, sp, ct, cap, var))
        descs[i] = QueryDescT(d_ir.ptr, len(ir), cap, int(var), int(depth), sp.ptr,
                              ct.ptr if ct else None)
    d_bounds = bounds.ptr if isinstance(bounds, DeviceBuffer) else int(bounds)
    d_m = DeviceBuffer.from_host(np.array([num_bounds], dtype=np.int64), stream)
    check(lib().gpj_query_run_records(ptr, int(n), ctypes.byref(index.raw), d_bounds, d_m.ptr,
                                      descs, len(compiled), records, stream), "query records")
    check(lib().gpj_stream_sync(stream))
    out = []
    for d_ir, sp, ct, cap, var in keep:
        spans = sp.to_host(np.int64, records * cap * 2, stream).reshape(records, cap, 2)
        counts = None
        if var:
            counts = ct.to_host(np.int64, records, stream)
            spans[np.arange(cap)[None, :] >= counts[:, None]] = -1
        out.append((spans, counts))
    return out


class narrow_span:
    """Context manager: the query kernels hand records reaching `nbytes` from
    their start (or backward searches reaching that far below it) to their
    64-bit kernel (default 2^31 - 2^20).  Results never depend on it; tests
    lower it to drive the 64-bit kernel on small inputs."""

    def __init__(self, nbytes: int):
        self.nbytes = nbytes

    def __enter__(self):
        check(lib().gpj_debug_set_narrow_span(self.nbytes), "narrow span")
        return self

    def __exit__(self, *exc):
        check(lib().gpj_debug_set_narrow_span(0), "narrow span")
        return False


class index_pairing:
    """Context manager: single-input index builds of at least `min_tiles`
    16 KiB tiles run two tiles per workgroup (default 8192; 1 = every input,
    PAIRING_NEVER = none).  Results never depend on it; tests drive both
    forms on small inputs."""

    def __init__(self, min_tiles: int):
        self.min_tiles = min_tiles

    def __enter__(self):
        check(lib().gpj_debug_set_index_pairing(self.min_tiles), "index pairing")
        return self

    def __exit__(self, *exc):
        check(lib().gpj_debug_set_index_pairing(0), "index pairing")
        return False


PAIRING_NEVER = 1 << 62

WALK_AUTO, WALK_LANE, WALK_WAVE = 0, 1, 2
WAVE_WALK_RECORDS, WAVE_WALK_LONG_RECORDS, WAVE_WALK_LONG_BYTES = 4096, 24576, 2048


def auto_walk_is_wave(n: int, records: int) -> bool:
    """The walk GPJ_WALK_AUTO picks for a 16-byte aligned input of n bytes and
    a record bound `records` (True: one record per wave)."""
    return records <= WAVE_WALK_RECORDS or (
        records <= WAVE_WALK_LONG_RECORDS and n >= WAVE_WALK_LONG_BYTES * records)


class walk:
    """Context manager: which query walk the query calls use -- WALK_LANE
    (one record per lane), WALK_WAVE (one record per wave, staged in LDS) or
    WALK_AUTO (the default, see auto_walk_is_wave).  Results never depend on
    it."""

    def __init__(self, mode: int):
        self.mode = mode

    def __enter__(self):
        check(lib().gpj_query_set_walk(self.mode), "query walk")
        return self

    def __exit__(self, *exc):
        check(lib().gpj_query_set_walk(WALK_AUTO), "query walk")
        return False


def debug_escape_quote(dev: DeviceBuffer | int, n: int, stream=None):
    ptr = dev.ptr if isinstance(dev, DeviceBuffer) else int(dev)
    w = (n + 63) // 64
    e, q = DeviceBuffer(max(w * 8, 8)), DeviceBuffer(max(w * 8, 8))
    check(lib().gpj_debug_escape_quote(ptr, int(n), e.ptr, q.ptr, stream), "escape/quote")
    check(lib().gpj_stream_sync(stream))
    return e.to_host(np.uint64, w), q.to_host(np.uint64, w)


# ---------------------------------------------------------------------------
# synthetic data (csrc/gpj_gen.c)
# ---------------------------------------------------------------------------
GEN_PATH = os.path.join(PKG_DIR, "_build", "libgpj_gen.so")
_gen = None

SCHEMA_NDJSON_2_4K = 0   # 64 MiB plain config: records of 2-4 KiB, query $.a.b
SCHEMA_BIG_ARRAY = 1     # one top-level array document
SCHEMA_NDJSON_1K = 2     # throughput config: ~1 KiB records, depth <= 4
SCHEMA_TWITTER = 3       # twitter-like records


def _gen_lib():
    global _gen
    if _gen is None:
        g = ctypes.CDLL(GEN_PATH)
        g.gpj_gen.argtypes = [ctypes.c_int, ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int,
                              ctypes.c_void_p, ctypes.c_uint64]
        g.gpj_gen.restype = ctypes.c_uint64
        g.gpj_gen_files.argtypes = [ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.c_uint64, ctypes.c_int, ctypes.c_uint64,
                                    ctypes.c_void_p, ctypes.c_void_p,
                                    ctypes.POINTER(ctypes.c_void_p)]
        g.gpj_gen_files.restype = ctypes.c_uint64
        g.gpj_gen_free.argtypes = [ctypes.c_void_p]
        _gen = g
    return _gen


def generate(schema: int, target: int, seed: int, threads: int | None = None) -> np.ndarray:
    """Seeded synthetic JSON as a uint8 array.  NDJSON schemas: `target` is the
    byte bound; schema 1: `target` is the element count."""
    _gen_lib()
    threads = threads or min(16, os.cpu_count() or 4)
    cap = target + (1 << 20) if schema != 1 else target * 1024 + (1 << 20)
    out = np.empty(cap, dtype=np.uint8)
    n = _gen.gpj_gen(schema, target, seed, threads, out.ctypes.data, cap)
    if n == 0:
        raise RuntimeError("generator buffer too small")
    return out[:n]


def generate_files(schemas, targets, seed: int, threads: int | None = None,
                   align: int = BATCH_ALIGN):
    """Many seeded files (NDJSON schemas 0, 2 or 3, each at most its target
    size and at least one record; schema 1: one top-level array document),
    laid out `align`-aligned in one zero-filled buffer.  Returns (buffer
    uint8, offsets uint64, sizes uint64)."""
    g = _gen_lib()
    sch = np.ascontiguousarray(schemas, dtype=np.int32)
    tg = np.ascontiguousarray(targets, dtype=np.uint64)
    f = sch.size
    offsets = np.zeros(f, dtype=np.uint64)
    sizes = np.zeros(f, dtype=np.uint64)
    ptr = ctypes.c_void_p()
    threads = threads or min(16, os.cpu_count() or 4)
    total = g.gpj_gen_files(f, sch.ctypes.data, tg.ctypes.data, seed, threads, align,
                            offsets.ctypes.data, sizes.ctypes.data, ctypes.byref(ptr))
    if total == 0:
        raise RuntimeError("gpj_gen_files failed")
    buf = np.empty(total, dtype=np.uint8)
    ctypes.memmove(buf.ctypes.data, ptr.value, total)
    g.gpj_gen_free(ptr)
    return buf, offsets, sizes


class FileBatch:
    """Many independent NDJSON files in one device buffer, indexed in one
    launch and queried in one launch (gpj_batch_index + gpj_batch_query,
    optionally captured in a HIP graph), or run with plain per-file launches
    (gpj_batch_per_file) for comparison.  Results of both land in the same
    layout: per slot k, spans [total lines][capacity][2] and counts [total
    lines]; file f's lines are line_prefix[f] .. line_prefix[f + 1]."""

    def __init__(self, buf: np.ndarray, offsets, sizes, masks, compiled, capacity: int = 8,
                 stream=None):
        L = lib()
        self.stream = stream
        self.n = int(buf.size)
        self.num_files = f = len(offsets)
        self.offsets = np.asarray(offsets, dtype=np.uint64)
        self.sizes = np.asarray(sizes, dtype=np.uint64)
        self.masks = np.asarray(masks, dtype=np.uint32)
        self.compiled = compiled
        self.capacity = capacity
        self.levels = max(1, max(c[1] for c in compiled))
        files = (BatchFileT * f)()
        for i in range(f):
            files[i] = BatchFileT(int(self.offsets[i]), int(self.sizes[i]), 0, 1,
                                  int(self.masks[i]), 0)
        self.words = (self.n + 63) // 64
        tiles = (self.n + BATCH_ALIGN - 1) // BATCH_ALIGN
        tile_file = np.zeros(max(tiles, 1), dtype=np.uint32)
        check(L.gpj_batch_tile_map(files, f, self.n, tile_file.ctypes.data), "tile map")
        self.d_in = DeviceBuffer.from_host(buf, stream)
        self.d_files = DeviceBuffer.from_host(
            np.frombuffer(bytes(files), dtype=np.uint8), stream)
        self.d_tile_file = DeviceBuffer.from_host(tile_file, stream)
        self.d_prefix = DeviceBuffer((f + 1) * 8)
        self.d_scratch = DeviceBuffer(L.gpj_batch_scratch_bytes(self.n))
        # each file's lines, counted on the device; its newline slice holds
        # exactly that many entries
        check(L.gpj_batch_line_counts(self.d_in.ptr, self.n, self.d_files.ptr, f,
                                      self.d_tile_file.ptr, self.d_prefix.ptr,
                                      self.d_scratch.ptr, stream), "batch line counts")
        check(L.gpj_stream_sync(stream))
        self.line_prefix = self.d_prefix.to_host(np.int64, f + 1, stream)
        self.line_counts = np.diff(self.line_prefix)
        self.total_lines = int(self.line_prefix[-1])
        for i in range(f):
            files[i].newline_offset = int(self.line_prefix[i])
            files[i].newline_capacity = int(self.line_counts[i])
        self.files = files
        self.d_files = DeviceBuffer.from_host(
            np.frombuffer(bytes(files), dtype=np.uint8), stream)
        self.d_string = DeviceBuffer(self.words * 8)
        self.d_leveled = DeviceBuffer(self.words * 8 * self.levels)
        self.d_newline = DeviceBuffer(self.total_lines * 8)
        self.descs = (QueryDescT * len(compiled))()
        self._irs = []
        for i, (ir, depth, slots) in enumerate(compiled):
            if slots > capacity:
                raise ValueError("result capacity below a query's slot count")
            d_ir = DeviceBuffer.from_host(np.frombuffer(ir, dtype=np.uint8), stream)
            self._irs.append(d_ir)
            self.descs[i] = QueryDescT(d_ir.ptr, len(ir), capacity,
                                       int(uses_extension_opcodes(ir)), int(depth), None, None)
        self.results = self._alloc_results()
        self.grid_records = self.total_lines
        self.per_wave = 64  # records per one-wave workgroup of the query launch
        self.graph = None

    def set_levels(self, levels: int):
        """Leveled rows the batch index builds (at least the queries'
        deepest level; more rows give the same results)."""
        need = max(1, max(c[1] for c in self.compiled))
        if levels < need:
            raise ValueError(f"the queries need {need} leveled rows")
        if levels != self.levels:
            self.release_graph()
            self.levels = levels
            self.d_leveled = DeviceBuffer(self.words * 8 * levels)

    def _alloc_results(self):
        sp = [DeviceBuffer(max(self.total_lines * self.capacity * 16, 16)) for _ in range(2)]
        ct = [DeviceBuffer(max(self.total_lines * 8, 8)) for _ in range(2)]
        r = BatchResultsT()
        for k in range(2):
            r.d_spans[k] = sp[k].ptr
            r.d_counts[k] = ct[k].ptr
        r.capacity = self.capacity
        return (r, sp, ct)

    def launch(self):
        """Batched index + query, stream-ordered (capturable)."""
        L = lib()
        check(L.gpj_batch_index(self.d_in.ptr, self.n, self.d_files.ptr, self.num_files,
                                self.d_tile_file.ptr, self.levels, self.d_string.ptr,
                                self.d_leveled.ptr, self.d_newline.ptr, self.d_prefix.ptr,
                                self.d_scratch.ptr, self.stream), "batch index")
        check(L.gpj_batch_query_shaped(self.d_in.ptr, self.n, self.d_files.ptr, self.num_files,
                                       self.levels, self.d_string.ptr, self.d_leveled.ptr,
                                       self.d_newline.ptr, self.d_prefix.ptr, self.grid_records,
                                       self.per_wave, self.descs, len(self.compiled),
                                       ctypes.byref(self.results[0]), self.stream),
              "batch query")

    def capture(self):
        """Capture launch() in a HIP graph (needs a non-default stream)."""
        L = lib()
        self.release_graph()
        check(L.gpj_graph_begin(self.stream), "graph begin")
        try:
            self.launch()
        finally:
            g = ctypes.c_void_p()
            st = L.gpj_graph_end(self.stream, ctypes.byref(g))
        check(st, "graph end")
        self.graph = g.value

    def launch_graph(self):
        check(lib().gpj_graph_launch(self.graph, self.stream), "graph launch")

    def release_graph(self):
        if self.graph:
            lib().gpj_graph_destroy(self.graph)
            self.graph = None

    def per_file_setup(self):
        L = lib()
        total_file_words = int(sum((int(z) + 63) // 64 for z in self.sizes))
        self.pf_string = DeviceBuffer(self.words * 8)
        self.pf_leveled = DeviceBuffer(max(total_file_words * 8 * self.levels, 8))
        self.pf_newline = DeviceBuffer(self.total_lines * 8)
        self.pf_counts = DeviceBuffer(self.num_files * 8)
        self.pf_scratch = DeviceBuffer(L.gpj_index_scratch_bytes(int(self.sizes.max())))
        self.pf_results = self._alloc_results()

    def launch_per_file(self):
        """Plain per-file launches (one index + one query launch per file)."""
        check(lib().gpj_batch_per_file(self.d_in.ptr, self.files, self.num_files, self.levels,
                                       self.pf_string.ptr, self.pf_leveled.ptr,
                                       self.pf_newline.ptr, self.pf_counts.ptr,
                                       self.line_prefix.ctypes.data, self.descs,
                                       len(self.compiled), ctypes.byref(self.pf_results[0]),
                                       self.pf_scratch.ptr, self.stream), "per-file launches")

    def check_status(self, per_file: bool = False):
        """Raises DeviceError when the last index launch failed (a tile's
        bounded wait for its carry gave up): the batched launch's flag, or a
        per-file line count of -1."""
        if per_file:
            check(lib().gpj_stream_sync(self.stream))
            if (self.pf_counts.to_host(np.int64, self.num_files, self.stream) < 0).any():
                raise DeviceError("per-file index build: a tile's wait for its carry timed out")
        else:
            index_status(self.d_scratch, self.n, self.stream)

    def fetch(self, per_file: bool = False):
        """Host copies: (slot spans list, slot counts list), entries past a
        record's count set to -1."""
        self.check_status(per_file)
        r = self.pf_results if per_file else self.results
        out_sp, out_ct = [], []
        for k in range(2):
            sp = r[1][k].to_host(np.int64, self.total_lines * self.capacity * 2, self.stream)
            sp = sp.reshape(self.total_lines, self.capacity, 2)
            ct = r[2][k].to_host(np.int64, self.total_lines, self.stream)
            sp[np.arange(self.capacity)[None, :] >= np.minimum(ct, self.capacity)[:, None]] = -1
            out_sp.append(sp)
            out_ct.append(ct)
        return out_sp, out_ct

    def device_line_prefix(self) -> np.ndarray:
        self.check_status()
        return self.d_prefix.to_host(np.int64, self.num_files + 1, self.stream)


def autotune_batch(batch: FileBatch, time_fn, cache_path: str | None = None):
    """Pick the launches of a file batch by measurement: the query launch's
    size (one grid pass over all records, or 4..32 waves per CU on 256 CUs
    looping over the records) and workgroup shape (64, 32 or 16 records per
    one-wave workgroup), then the number of leveled rows (the queries'
    deepest level or one more).  The index tile (16 KiB, one 256-lane
    workgroup) is fixed by the batch layout: files start on tiles.
    `time_fn(batch)` returns the time of one batched run in ms.  The choice
    is cached by (record-count bucket, levels, query count) in cache_path
    (JSON) when given."""
    import json as _json
    key = (f"lines~2^{max(batch.total_lines, 1).bit_length()}/levels{batch.levels}"
           f"/q{len(batch.compiled)}/shape")
    cache = {}
    if cache_path and os.path.exists(cache_path):
        with open(cache_path) as fh:
            cache = _json.load(fh)
    if key in cache:
        batch.grid_records = int(cache[key]["grid_records"])
        batch.per_wave = int(cache[key]["per_wave"])
        batch.set_levels(int(cache[key].get("levels", batch.levels)))
        return cache[key]
    trials = []
    for per_wave in (64, 32, 16):
        waves = [None] + [256 * w for w in (4, 8, 16, 32)]  # None: one pass over all records
        for w in waves:
            g = batch.total_lines if w is None else w * per_wave
            batch.grid_records = max(per_wave, min(g, batch.total_lines))
            batch.per_wave = per_wave
            trials.append((time_fn(batch), batch.grid_records, per_wave))
    best = min(trials)
    batch.grid_records, batch.per_wave = best[1], best[2]
    # leveled rows: the queries' deepest level, or one more
    base = batch.levels
    level_trials = [(best[0], base)]
    batch.set_levels(base + 1)
    level_trials.append((time_fn(batch), base + 1))
    best_levels = min(level_trials)[1]
    batch.set_levels(best_levels)
    choice = {"grid_records": best[1], "per_wave": best[2], "ms": round(best[0], 4),
              "trials": [{"grid_records": g, "per_wave": p, "ms": round(t, 4)}
                         for t, g, p in trials],
              "levels": best_levels,
              "level_trials": [{"levels": lv, "ms": round(t, 4)} for t, lv in level_trials],
              "index_tile_bytes": BATCH_ALIGN,
              "index_tile_note": "fixed: files start on index tiles (GPJ_BATCH_ALIGN), one "
                                 "workgroup of 256 lanes per tile"}
    if cache_path:
        cache[key] = choice
        with open(cache_path, "w") as fh:
            _json.dump(cache, fh, indent=1)
    return choice

