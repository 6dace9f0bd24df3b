"""The PDHG benchmark runner (vikalp_pdhg_benchmark), the counterpart of the
reference's tools/benchmark_runner: its command line on the CPU, and one
case solved on the GPU with its CSV and JSON records checked against the
oracle and HiGHS."""
import csv
import json
import os
import subprocess

import pytest

import vikalp_amd as V
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNNER = os.path.join(ROOT, "vikalp_amd", "lib", "vikalp_pdhg_benchmark")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def run(*args, timeout=300):
    return subprocess.run([RUNNER, *args], capture_output=True, text=True, timeout=timeout)


def test_runner_command_line():
    assert os.path.exists(RUNNER), "build the package first (python -m vikalp_amd.build)"
    r = run("--help")
    assert r.returncode == 0 and "--csv" in r.stdout and "zipf_400000x1000000" in r.stdout
    r = run("--backend", "cpu")
    assert r.returncode == 2 and "HIP backend only" in r.stderr
    assert run("--policy", "simplex").returncode == 2
    assert run("--unknown").returncode == 2


@pytest.mark.gpu
def test_runner_records_on_the_config1_lp(tmp_path):
    csv_path, json_path = tmp_path / "r.csv", tmp_path / "r.json"
    r = run("--cases", "uniform_1000x3000", "--csv", str(csv_path), "--json", str(json_path))
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    rows = list(csv.DictReader(open(csv_path)))
    recs = json.load(open(json_path))
    assert len(rows) == 1 and len(recs) == 1
    row, rec = rows[0], recs[0]
    assert row["case"] == rec["case"] == "uniform_1000x3000"
    assert rec["status"] == "Optimal" and rec["variables"] == 3000 and rec["constraints"] == 1000
    assert rec["relative_kkt"] <= 1e-6 * 1.0001
    highs = json.load(open(os.path.join(GOLDEN, "highs_optima.json")))["uniform_1000x3000_k5"]["objective"]
    assert abs(rec["objective"] - highs) <= 1e-5 * max(1.0, abs(highs))
    assert float(row["objective"]) == rec["objective"]


@pytest.mark.gpu
def test_runner_reference_policy_matches_oracle(tmp_path):
    json_path = tmp_path / "r.json"
    r = run("--cases", "uniform_1000x3000", "--policy", "reference", "--iteration-limit", "2000",
            "--json", str(json_path))
    assert r.returncode == 0, r.stderr
    rec = json.load(open(json_path))[0]
    ref = O.solve_lp(V.generate_lp(1, 1000, 3000, 5, 20261203).as_oracle(), iteration_limit=2000)
    assert rec["status"] == ref.status and rec["iterations"] == ref.iterations
    assert abs(rec["objective"] - ref.objective_value) <= 1e-9 * max(1.0, abs(ref.objective_value))
