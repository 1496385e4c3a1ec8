"""GPU checks of the request server (gpjson::serve::Server through the
gpjson_serve replay tool).  Every served request, at every arrival rate, must
equal the same request served alone (the tool checks that) and the CPU
oracle run on its document: the tool writes the document manifest and a
digest of each request's results (CSR offsets and spans per query); the test
regenerates the documents, runs the oracle once per (document, query) and
compares every digest.  The configured serving workload: Poisson arrivals at
500, 1000, 2000 and 5000 requests/s, documents of 4 KiB to 64 MiB
(log-uniform) from the three record schemas, 1 to 8 queries each from that
schema's configuration, 30% re-queries of resident documents, an 80 GiB HBM
budget."""
import json
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "15418-project_amd", "_build", "gpjson_serve")
THREADS = 16


def run_serve(*args, timeout=540):
    print(f"[serve] {' '.join(args)}", flush=True)
    p = subprocess.run([SERVE, *args], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr[-2000:]
    out = [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]
    for o in out:
        print(json.dumps(o), flush=True)
    return out


_C0, _C1, _C2 = (np.uint64(0x9E3779B97F4A7C15), np.uint64(0xBF58476D1CE4E5B9),
                 np.uint64(0x94D049BB133111EB))


def digest(values) -> int:
    """The tool's digest: sum of value * (SplitMix64(index) | 1) mod 2^64."""
    v = np.ascontiguousarray(values, dtype=np.int64).ravel().view(np.uint64)
    z = np.arange(v.size, dtype=np.uint64) + _C0
    z = (z ^ (z >> np.uint64(30))) * _C1
    z = (z ^ (z >> np.uint64(27))) * _C2
    w = (z ^ (z >> np.uint64(31))) | np.uint64(1)
    return int(np.sum(v * w, dtype=np.uint64))


def oracle_results(gpj, orc, doc: bytes, levels: int, text: str, compiled):
    """(records, offsets, spans) of one query over one document, from the
    oracle: lines of an NDJSON document, or the elements of a top-level array
    for `$[*]...` queries."""
    ir, _depth, slots = compiled
    variable = gpj.uses_extension_opcodes(ir)
    ix = orc.build_index(doc, levels, THREADS)
    if text.startswith("$[*]"):
        bounds = orc.row_bounds(ix)
        cap = 32
        while True:
            sp, ct = orc.query_records(doc, ix, bounds, ir, cap if variable else slots, variable)
            if not variable or int(ct.max(initial=0)) <= cap:
                break
            cap = int(ct.max())
        off, spans = gpj.dense_from_slots(sp, ct if variable else None)
        return max(bounds.size - 1, 0), off, spans
    if variable:
        cap = 8
        while True:
            sp, ct = orc.query_ext(doc, ix, ir, cap, THREADS)
            if int(ct.max(initial=0)) <= cap:
                break
            cap = int(ct.max())
        off, spans = gpj.dense_from_slots(sp, ct)
    else:
        off, spans = gpj.dense_from_slots(orc.query(doc, ix, ir, slots, THREADS))
    return ix[2].size, off, spans


def check_digests(gpj, orc, path: str) -> int:
    with open(path) as f:
        lines = [json.loads(x) for x in f]
    man = lines[0]["manifest"]
    buf, off, sz = gpj.generate_files(np.array(man["schemas"], dtype=np.int32),
                                      np.array(man["targets"], dtype=np.uint64), man["seed"],
                                      THREADS, align=man["align"])
    np.testing.assert_array_equal(off, man["offsets"])
    np.testing.assert_array_equal(sz, man["sizes"])
    compiled = [gpj.compile_element_query(t) if t.startswith("$[*]")
                else gpj.compile_query(t, extensions=True) for t in man["queries"]]
    cache, checked = {}, 0
    for rec in lines[1:]:
        d = rec["doc"]
        for k, qi in enumerate(rec["queries"]):
            key = (d, qi)
            if key not in cache:
                doc = bytes(buf[int(off[d]):int(off[d] + sz[d])])
                n, o, s = oracle_results(gpj, orc, doc, man["levels"], man["queries"][qi],
                                         compiled[qi])
                cache[key] = (n, digest(o), digest(s))
            n, do, ds = cache[key]
            assert rec["records"] == n, (rec, n)
            assert rec["digests"][2 * k] == do, (rec["rate"], rec["request"], qi, "offsets")
            assert rec["digests"][2 * k + 1] == ds, (rec["rate"], rec["request"], qi, "spans")
        checked += 1
    print(f"[serve] {checked} request results equal the oracle "
          f"({len(cache)} document/query pairs)", flush=True)
    return checked


@pytest.mark.timeout(900)
def test_configured_workload_matches_oracle(gpu, orc, tmp_path):
    path = str(tmp_path / "digests.jsonl")
    out = run_serve("--requests", "1000", "--rates", "500,1000,2000,5000", "--min-doc", "4KiB",
                    "--max-doc", "64MiB", "--requery", "0.3", "--budget-gib", "80",
                    "--slots", "8", "--verify", "-1", "--seed", "0x15418004",
                    "--digest-out", path)
    assert [o["rate_per_s"] for o in out] == [500, 1000, 2000, 5000]
    for o in out:
        assert o["identical"] is True and o["verified"] == 1000
        assert o["resident_hits"] > 0
        assert o["p99_ms"] >= o["p50_ms"] > 0
        assert o["peak_device_bytes"] <= o["hbm_budget_bytes"]
    assert check_digests(gpu, orc, path) == 4000


def test_mixed_requests_equal_solo_runs(gpu, orc, tmp_path):
    path = str(tmp_path / "digests.jsonl")
    out = run_serve("--requests", "150", "--rates", "400,4000", "--max-doc", "2MiB",
                    "--slots", "6", "--verify", "-1", "--seed", "11", "--digest-out", path)
    assert len(out) == 2
    for o in out:
        assert o["identical"] is True and o["verified"] == 150
        assert o["resident_hits"] > 0
    assert check_digests(gpu, orc, path) == 300


def test_small_budget_evicts_and_stays_exact(gpu, orc, tmp_path):
    path = str(tmp_path / "digests.jsonl")
    out = run_serve("--requests", "120", "--rates", "3000", "--max-doc", "1MiB",
                    "--budget-gib", "0.02", "--slots", "4", "--verify", "-1", "--seed", "12",
                    "--digest-out", path)
    o = out[0]
    assert o["identical"] is True and o["verified"] == 120
    assert o["evictions"] > 0
    assert o["peak_device_bytes"] <= o["hbm_budget_bytes"]
    assert check_digests(gpu, orc, path) == 120


def test_capacity_guesses_rerun_exactly(gpu, orc, tmp_path):
    # guesses far too small for the lines, the array elements and the spans:
    # every kind of rerun happens, and the results stay exact
    path = str(tmp_path / "digests.jsonl")
    out = run_serve("--requests", "90", "--rates", "2000", "--min-doc", "64KiB",
                    "--max-doc", "4MiB", "--bytes-per-line", "16384",
                    "--bytes-per-element", "8192", "--spans-per-record", "1", "--slots", "4",
                    "--verify", "-1", "--seed", "13", "--digest-out", path)
    o = out[0]
    assert o["identical"] is True and o["verified"] == 90
    assert o["retries"] > 0
    assert check_digests(gpu, orc, path) == 90
