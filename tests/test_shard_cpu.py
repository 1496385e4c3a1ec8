"""Host side of indexing one document in shards: gpj_shard_carries composes
the shards' summaries into each shard's carry-in.  Checked against the
sequential state scan of the whole document at every cut (tests/shardstate.py)
on inputs dense in backslash runs, quotes, brackets and newlines, with cuts
inside strings, inside backslash runs and at depth."""
import random

import pytest

import jsongen
from shardstate import scan, summary


def carries_of(gpj, data: bytes, cuts):
    bounds = [0, *cuts, len(data)]
    sums = [gpj.ShardSummaryT(*summary(data[a:b])) for a, b in zip(bounds, bounds[1:])]
    return gpj.shard_carries(sums)


@pytest.mark.parametrize("seed", range(6))
def test_carries_equal_the_sequential_scan(gpj, seed):
    r = random.Random(seed)
    if seed % 2:
        data = jsongen.byte_soup(seed, 20_000, special_weight=0.5)
    else:
        data = jsongen.ndjson(seed, 20_000, max_depth=7)
    cuts = sorted(r.sample(range(64, len(data) - 1, 64), 12))
    carries = carries_of(gpj, data, cuts)
    assert len(carries) == len(cuts) + 2
    c0 = carries[0]
    assert (c0.escape, c0.in_string, c0.depth, c0.newlines) == (0, 0, -1, 0)
    for c, cut in zip(carries[1:], cuts + [len(data)]):
        assert (c.escape, c.in_string, c.depth, c.newlines) == scan(data[:cut]), cut


def test_cuts_inside_backslash_runs_and_strings(gpj):
    # runs of 1..9 backslashes inside and outside strings, cut after every byte
    data = b'{"a":"x\\\\\\"y","b":[' + b'\\' * 9 + b'"' + b'\\' * 4 + b'{[' * 40 + b'"z' * 7 + b'\n' * 3
    for cut in range(1, len(data)):
        carries = carries_of(gpj, data, [cut])
        c = carries[1]
        assert (c.escape, c.in_string, c.depth, c.newlines) == scan(data[:cut]), cut
        e = carries[2]
        assert (e.escape, e.in_string, e.depth, e.newlines) == scan(data), cut


def test_depth_wraps_like_int8(gpj):
    data = b"[" * 300 + b"]" * 10
    carries = carries_of(gpj, data, [128, 256])
    assert [c.depth for c in carries] == [scan(data[:k])[2] for k in (0, 128, 256, len(data))]
    assert carries[-1].depth == ((-1 + 290 + 128) % 256) - 128


@pytest.mark.parametrize("seed", [3, 4])
def test_sequential_scan_agrees_with_the_oracle_index(orc, seed):
    # the scan's in-string state after each byte is the oracle's string mask
    # bit there, and every structural byte outside strings sits in the
    # oracle's leveled row of its depth (an opener at the depth it opens)
    import numpy as np
    data = (jsongen.byte_soup(seed, 6_000, special_weight=0.5) if seed % 2
            else jsongen.ndjson(seed, 6_000, max_depth=6))
    levels = 8
    string, leveled, _ = orc.build_index(data, levels)
    smask = np.unpackbits(np.asarray(string).view(np.uint8), bitorder="little")
    rows = np.unpackbits(np.asarray(leveled).reshape(levels, -1).view(np.uint8), axis=1,
                         bitorder="little")
    esc, instr, depth = 0, 0, -1
    for i, b in enumerate(data):
        before_depth = depth
        esc2, instr, depth, _ = scan(bytes([b]), esc, instr, depth, 0)
        assert smask[i] == instr, i
        structural = not instr and b in b"{}[]:," and not (b == 0x22)
        if b in b"{[" and structural:
            want = depth
        elif structural:
            want = before_depth
        else:
            want = None
        for k in range(levels):
            assert rows[k][i] == (1 if want == k else 0), (i, k, want)
        esc = esc2
