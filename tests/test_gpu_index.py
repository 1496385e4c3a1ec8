"""GPU parity of the structural index: every string-mask word, leveled word
and newline offset must equal the CPU oracle bit for bit (integer work)."""
import numpy as np
import pytest

import jsongen

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["single", "pairs"])
def pairing(request, gpu):
    # every test under both index forms: one tile per workgroup and two
    # (the default picks two from 128 MiB on)
    with gpu.index_pairing(1 if request.param == "pairs" else gpu.PAIRING_NEVER):
        yield request.param

FIXTURE = (b'{"id":1,"name":"Ada","tags":["gpu","json"],"active":true}\n'
           b'{"id":2,"name":"Bob \\"The Builder\\"","note":"literal { braces } and [brackets] stay strings"}\n'
           b'{"id":3,"nested":{"arr":[1,{"x":"y,z"}],"empty":{}},"path":"C:\\\\tmp\\\\file"}')


def check_index(gpj, orc, data: bytes, levels: int, newline_capacity: int = 0):
    print(f"[index] n={len(data)} levels={levels} cap={newline_capacity}", flush=True)
    s, lv, nl = orc.build_index(data, levels)
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    idx = gpj.build_index(dev, len(data), levels, newline_capacity=newline_capacity)
    assert idx.num_lines == nl.size
    np.testing.assert_array_equal(idx.newline(), nl)
    got_s = idx.string()
    if not np.array_equal(got_s, s):
        bad = np.nonzero(got_s != s)[0]
        raise AssertionError(f"string mask differs at {bad.size} words, first {bad[:8]}")
    got_l = idx.leveled()
    if not np.array_equal(got_l, lv):
        bad = np.argwhere(got_l != lv)
        raise AssertionError(f"leveled differs at {len(bad)} words, first {bad[:8].tolist()}")
    idx.free()
    dev.free()


def test_reference_fixture_all_levels(gpu, orc):
    # tests/index/index_test_utils.hpp:80-89 fixture, reference kMaxDepth = 8
    for levels in (1, 2, 3, 8, 16, 22):
        check_index(gpu, orc, FIXTURE, levels)


def test_reference_kats(gpu, orc):
    check_index(gpu, orc, b'"a\\"b"  "x"', 1)
    check_index(gpu, orc, b"first\nsecond\nthird", 1)
    check_index(gpu, orc, b'{"a":[1,2]}', 3)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 127, 128, 1000, 16383, 16384, 16385,
                               32768, 40000, 100003])
def test_sizes_byte_soup(gpu, orc, n):
    check_index(gpu, orc, jsongen.byte_soup(n, n), 4)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_ndjson_multi_tile(gpu, orc, seed):
    data = jsongen.ndjson(seed, 3 << 20, max_depth=7, spaces=(seed == 2))
    check_index(gpu, orc, data, 8)


def test_empty_input(gpu, orc):
    dev = gpu.DeviceBuffer(16)
    idx = gpu.build_index(dev, 0, 3)
    assert idx.num_lines == 1
    np.testing.assert_array_equal(idx.newline(), np.array([0], dtype=np.int64))
    assert idx.level_size == 0


def test_backslash_runs_across_tiles(gpu, orc):
    # whole 16 KiB tiles of backslashes, odd and even runs ending on quotes
    parts = [b'{"a":"', b"\\" * 40001, b'"', b'",":{[', b"\\" * 16384, b'"x"', b"\\" * 16383,
             b'"{}[],:"', b"\\" * 65536, b'"\n', b'{"b":1}']
    check_index(gpu, orc, b"".join(parts), 4)
    # a run ending exactly on a tile boundary, quote as the first byte of a tile
    data = b"x" * (16384 - 7) + b"\\" * 7 + b'"{"a":1}"' + b"[" * 50
    check_index(gpu, orc, data, 4)
    data = b"x" * (16384 - 6) + b"\\" * 6 + b'"{"a":1}"' + b"[" * 50
    check_index(gpu, orc, data, 4)


def soup_np(seed: int, n: int) -> bytes:
    # bytes dense in the structural characters, with runs of backslashes
    rng = np.random.default_rng(seed)
    palette = np.frombuffer(b'\\"{}[]:,\nabcd x', dtype=np.uint8)
    out = palette[rng.integers(0, palette.size, n)]
    for at in rng.integers(0, n - 200, n // 20000):
        out[at:at + int(rng.integers(1, 140))] = ord("\\")
    return out.tobytes()


def test_many_scan_rounds(gpu, orc):
    # more 16 KiB tiles than one round of the carry scan resolves (1,024),
    # with escape, string and depth carries across every tile, every group
    # of 8 tiles and every round; one string spans more than a round
    long_string = b'"' + b"x" * (17 << 20) + b'"'
    data = soup_np(7, 20 << 20) + long_string + soup_np(8, (20 << 20) + 12345)
    check_index(gpu, orc, data, 4)


def test_strings_spanning_tiles(gpu, orc):
    long_string = b'"' + b"{[:,]}" * 20000 + b'"'
    data = b'{"a":' + long_string + b',"b":[' + long_string + b',{"c":2}]}\n' * 3
    check_index(gpu, orc, data, 4)


def test_deep_nesting_wraps_like_int8(gpu, orc):
    # the reference keeps depth in a signed char: 300 levels wrap modulo 256
    data = b"[" * 300 + b"1" + b"]" * 300 + b"\n" + b"{" * 129 + b"}" * 129
    check_index(gpu, orc, data, 22)
    data = (b'{"a":' * 140 + b"0" + b"}" * 140 + b"\n") * 100
    check_index(gpu, orc, data, 22)


def test_non_ascii_aliases(gpu, orc):
    aliases = bytes(c | 0x80 for c in b'\\"{}[]:,\n')
    data = (b'{"k":"' + aliases + b'","\xa2":[\xdb\xdd]}\n') * 2000
    check_index(gpu, orc, data, 3)


def test_newline_capacity_fallback(gpu, orc):
    data = b"\n" * 50000 + b'{"a":1}\n' * 10000
    check_index(gpu, orc, data, 2, newline_capacity=10)


def test_debug_escape_quote(gpu, orc):
    for seed, n in ((5, 5000), (6, 70001)):
        data = jsongen.byte_soup(seed, n)
        dev = gpu.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
        e, q = gpu.debug_escape_quote(dev, len(data))
        np.testing.assert_array_equal(e, orc.escape_index(data))
        np.testing.assert_array_equal(q, orc.quote_index(data))
