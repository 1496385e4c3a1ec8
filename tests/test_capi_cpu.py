"""CPU-only checks of the C ABI library: it loads, exports every function the
header declares, and its host-only entry points behave without a GPU."""
import ctypes

import pytest


def test_library_exports_every_declared_symbol(gpj):
    L = gpj.lib()
    names = gpj.exported_symbols_declared_in_header()
    assert len(names) >= 15, names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_compile_needs_no_device(gpj):
    ir, depth, slots = gpj.compile_query("$.name")
    assert ir == b"\x04\x04name\x03\x01\x00" and depth == 1 and slots == 1


def test_compile_errors_map_to_reference_classes(gpj):
    with pytest.raises(gpj.MalformedQueryError):
        gpj.compile_query("$.")
    with pytest.raises(gpj.MalformedQueryError):
        gpj.compile_query("$.arr[-0]")
    with pytest.raises(gpj.UnsupportedQueryError):
        gpj.compile_query("$..name")
    with pytest.raises(gpj.UnsupportedQueryError):
        gpj.compile_query("$.arr[*]")


def test_bad_arguments_rejected_before_any_launch(gpj):
    L = gpj.lib()
    raw = gpj.IndexT()
    # levels above the reference's 22-level limit (leveled-bitmaps-index.cu:6)
    st = L.gpj_index_build(None, 0, 23, None, ctypes.byref(raw), None)
    assert st == gpj.ERR_INVALID_ARGUMENT
    assert b"levels" in L.gpj_last_error()
    # IR longer than the reference's 4096-byte buffer (query_executor_host.cu:190-192)
    st = L.gpj_query_run(None, 10, ctypes.byref(raw), None, 5000, 1, 1, None, None)
    assert st == gpj.ERR_QUERY_EXECUTION
