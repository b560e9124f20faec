"""bench.py's command-line contract on the CPU: one GPU only."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True,
                          text=True, timeout=120, env={**os.environ, **(env or {})})


def test_more_than_one_gpu_exits_with_a_message():
    r = _run(["--gpus", "2"])
    assert r.returncode != 0
    assert "one GPU" in r.stderr and r.stdout == ""


def test_launch_with_world_size_above_one_exits_with_a_message():
    r = _run(["--gpus", "1"], env={"WORLD_SIZE": "4"})
    assert r.returncode != 0 and "one GPU" in r.stderr


def test_secondary_summary_keeps_the_figures():
    sys.path.insert(0, ROOT)
    import bench
    rec = {"metric": "m", "value": 1.5, "unit": "TFLOP/s", "ms_per_step": 2.0, "dtype": "f64",
           "config": {"workload": "w"}, "roofline": {"kernel": "k", "achieved": 1.0, "peak": 2.0, "unit": "TFLOP/s",
                                                      "frac": 0.5, "traffic": None, "extra": 1},
           "cpu_baseline": {"value": 0.1, "unit": "TFLOP/s", "cores": 16, "kind": "port", "sample": "s"},
           "capacity": {"matvecs_per_s": 10.0}, "latency_ms": {"p50": 0.1, "p99": 0.2},
           "cells": [{"symmetry": "U1", "d": 4, "workload": "compose", "us_per_iter": 7.123,
                      "us_per_call_c_abi": 5.6, "reference_us": 4.1, "checksum": 61.125}]}
    out = bench._summary(rec)
    assert out["workload"] == "w" and out["value"] == 1.5 and out["roofline"]["frac"] == 0.5
    assert "extra" not in out["roofline"] and out["cpu_baseline"]["cores"] == 16
    assert out["capacity_matvecs_per_s"] == 10.0 and out["latency_ms"]["p99"] == 0.2
    assert out["cells"] == [{"case": "U1 d=4 compose", "us": 7.12, "us_c_abi": 5.6, "reference_us": 4.1,
                             "checksum": 61.125}]
    names = [n for n, _, _ in bench.SECONDARY]
    assert len(names) == len(set(names)) and all(limit > 0 for _, _, limit in bench.SECONDARY)
