"""bench.py's command-line contract, checked on the CPU (no GPU work):
one GPU only, and the JSON line's documented keys."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args, env=None):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, env=env)


def test_more_than_one_gpu_is_refused():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_world_size_above_one_is_refused():
    env = dict(os.environ, WORLD_SIZE="4")
    r = _run(env=env)
    assert r.returncode == 2
    assert "WORLD_SIZE=4" in r.stderr


def test_unknown_workload_is_refused():
    r = _run("--workload", "nonsense")
    assert r.returncode != 0 and "invalid choice" in r.stderr
