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
