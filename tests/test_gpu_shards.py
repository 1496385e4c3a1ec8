"""GPU parity of indexing one document in shards (the share of one GPU when a
document is split over several): each shard is built alone and summarized,
the summaries compose into carry-ins on the host, and each shard rebuilt from
its carry-in must give exactly the whole document's string mask and leveled
rows from its first word on, and its newlines.  Cuts fall inside strings,
inside backslash runs and at depth; the summaries are also checked against the
sequential state scan (tests/shardstate.py)."""
import numpy as np
import pytest

import jsongen
from shardstate import scan

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True, params=["single", "pairs"])
def pairing(request, gpu):
    # every test under both index forms: one tile per workgroup and two
    # (the default picks two from 128 MiB on)
    with gpu.index_pairing(1 if request.param == "pairs" else gpu.PAIRING_NEVER):
        yield request.param


def check_shards(gpj, data: bytes, cuts, levels: int, label: str):
    n = len(data)
    dev = gpj.DeviceBuffer.from_host(np.frombuffer(data, dtype=np.uint8))
    whole = gpj.ShardIndex(dev, n, levels)
    w_str, w_lev, w_nl = whole.string_words(), whole.leveled_words(), whole.newlines()
    bounds = [0, *cuts, n]
    # pass 1: each shard alone, summarized
    sums = []
    for a, b in zip(bounds, bounds[1:]):
        sums.append(gpj.ShardIndex(dev.ptr + a, b - a, levels).summary())
    carries = gpj.shard_carries(sums)
    state = (0, 0, -1, 0)
    for c, a, b in zip(carries, [0, *bounds], bounds):
        state = scan(data[a:b], *state)
        assert (c.escape, c.in_string, c.depth, c.newlines) == state, (label, b)
    # pass 2: each shard from its carry-in
    for k, (a, b) in enumerate(zip(bounds, bounds[1:])):
        print(f"[shards] {label} shard {k} [{a}, {b}) carry=({carries[k].escape}, "
              f"{carries[k].in_string}, {carries[k].depth}, {carries[k].newlines})", flush=True)
        s = gpj.ShardIndex(dev.ptr + a, b - a, levels, carry=carries[k])
        w0, w = a // 64, s.words
        np.testing.assert_array_equal(s.string_words(), w_str[w0:w0 + w])
        np.testing.assert_array_equal(s.leveled_words(), w_lev[:, w0:w0 + w])
        nl = s.newlines()
        first = int(carries[k].newlines)
        count = int(carries[k + 1].newlines) - first
        assert nl.size == count + 1 and nl[0] == 0
        np.testing.assert_array_equal(nl[1:] + a, w_nl[first + 1:first + 1 + count])


@pytest.mark.parametrize("seed", [71, 72])
def test_ndjson_document_in_shards(gpu, seed):
    data = jsongen.ndjson(seed, 3 << 20, max_depth=7, spaces=(seed == 72))
    r = np.random.default_rng(seed)
    cuts = sorted(int(c) * 64 for c in r.choice(len(data) // 64, 7, replace=False) if c > 0)
    check_shards(gpu, data, cuts, 4, f"ndjson seed={seed}")


def test_byte_soup_in_shards(gpu):
    # dense quotes, backslash runs and brackets: cuts land in every state
    data = jsongen.byte_soup(73, 1 << 20, special_weight=0.5)
    cuts = list(range(64 * 1001, len(data), 64 * 4093))
    check_shards(gpu, data, cuts, 6, "byte soup")


def test_array_document_in_shards(gpu):
    # one top-level array (no newlines): every element carries depth >= 1
    data = bytes(gpu.generate(gpu.SCHEMA_BIG_ARRAY, 40_000, 0x15418001))
    n = len(data)
    cuts = [64 * (n // 64 // 5 * k) for k in range(1, 5)]
    check_shards(gpu, data, cuts, 8, "array")


def test_cuts_in_strings_and_escape_runs(gpu):
    # cuts right after odd and even backslash runs, inside a long string and
    # deep in nesting (positions chosen on the host from the scan)
    unit = (b'{"s":"' + b"x" * 50 + b'\\\\\\"' + b"y" * 9 + b'","d":' + b"[" * 20 + b"1" +
            b"]" * 20 + b',"e":"\\\\' + b'z' * 60 + b'"}\n')
    data = unit * 400
    states, cuts = {}, []
    esc = instr = 0
    depth, nl = -1, 0
    for i in range(64, len(data), 64):
        esc, instr, depth, nl = scan(data[i - 64:i], esc, instr, depth, nl)
        key = (esc, instr, depth > 0)
        if key not in states:
            states[key] = i
            cuts.append(i)
    assert len(states) >= 3
    check_shards(gpu, data, sorted(cuts), 4, "strings and escapes")


def test_cuts_inside_long_backslash_runs(gpu):
    # runs of 63..200 backslashes (whole words of them pass the escape carry
    # through), cut at every 64-byte multiple
    parts = []
    for run in (63, 64, 65, 127, 128, 129, 200):
        parts.append(b'{"s":"' + b"\\" * run + b'x","t":[' + b"\\" * (run // 2) + b']}\n')
    data = b"".join(parts) * 3
    cuts = list(range(64, len(data), 64))
    check_shards(gpu, data, cuts, 3, "backslash runs")
