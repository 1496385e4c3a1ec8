"""Runs the C++ API test binary (tests/cpp/test_gpjson.cpp): the reference's
gtest cases re-expressed against the gpjson:: host API."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BUILD = os.path.join(ROOT, "15418-project_amd", "_build")
TESTS_BIN = os.path.join(BUILD, "gpjson_cpp_tests")
DRIVER = os.path.join(BUILD, "gpjson_driver")


def run(args, timeout=300):
    return subprocess.run(args, capture_output=True, text=True, timeout=timeout)


def test_cpp_api_cpu_cases():
    r = run([TESTS_BIN, "--cpu"])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


def test_driver_cli_surface(tmp_path):
    r = run([DRIVER, "--help"])
    assert r.returncode == 0 and "--index-builder.type" in r.stdout and "--query-executor.grid-size" in r.stdout
    r = run([DRIVER, "--dataset", "x.json"])
    assert r.returncode == 1 and "At least one query is required" in r.stderr
    r = run([DRIVER, "--dataset", "x.json", "-q", "$.a", "--index-builder.type", "NOPE"])
    assert r.returncode == 1 and "Invalid enum value 'NOPE'" in r.stderr


@pytest.mark.gpu
def test_cpp_api_gpu_cases():
    r = run([TESTS_BIN, "--gpu"])
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr


@pytest.mark.gpu
def test_driver_end_to_end(tmp_path):
    data = tmp_path / "d.ldjson"
    data.write_bytes(b'{"name":"Ada","scores":[1,2],"lang":"en"}\n'
                     b'{"name":"Bob","scores":[3,4],"lang":"fr"}\n')
    cfg = tmp_path / "c.toml"
    cfg.write_text(f'dataset_path = "{data}"\nrun_name = "t"\nqueries = ["$.name", "$.scores[1]"]\n'
                   '[file_reader]\ntype = "MMAP"\n[index_builder]\ntype = "COMBINED"\n'
                   'file_partition_size = 43\n')
    r = run([DRIVER, "--config", str(cfg)])
    assert r.returncode == 0, r.stderr
    assert "### RUN BEGIN run_name=t ###" in r.stdout
    assert "Query '$.name' has 2 lines of result, and 2 matched values" in r.stdout
    assert "Query '$.scores[1]' has 2 lines of result, and 2 matched values" in r.stdout
    assert "### RUN END run_name=t ###" in r.stdout
