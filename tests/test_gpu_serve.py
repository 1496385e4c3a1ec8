"""GPU checks of the request server (gpjson::serve::Server through the
gpjson_serve replay tool): every request's results equal the same request
served alone, under overlap, residency and HBM-budget evictions."""
import json
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SERVE = os.path.join(ROOT, "15418-project_amd", "_build", "gpjson_serve")


def run_serve(*args):
    print(f"[serve] {' '.join(args)}", flush=True)
    p = subprocess.run([SERVE, *args], capture_output=True, text=True, timeout=240)
    assert p.returncode == 0, p.stderr[-2000:]
    return [json.loads(line) for line in p.stdout.splitlines() if line.startswith("{")]


def test_mixed_requests_equal_solo_runs(gpu):
    out = run_serve("--requests", "150", "--rates", "400,4000", "--max-doc", "2MiB",
                    "--slots", "6", "--verify", "-1", "--seed", "11")
    assert len(out) == 2
    for o in out:
        assert o["identical"] is True and o["verified"] > 0
        assert o["resident_hits"] > 0
        assert o["p99_ms"] >= o["p50_ms"] > 0


def test_small_budget_evicts_and_stays_exact(gpu):
    out = run_serve("--requests", "120", "--rates", "3000", "--max-doc", "1MiB",
                    "--budget-gib", "0.02", "--slots", "4", "--verify", "-1", "--seed", "12")
    o = out[0]
    assert o["identical"] is True and o["verified"] == 120
    assert o["evictions"] > 0
    assert o["peak_device_bytes"] <= o["hbm_budget_bytes"]
