"""bench.py's command-line contract on the CPU: more than one GPU is refused with a message
and exit code 2 before anything touches a GPU; invalid in-flight depths are rejected."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_more_than_one_gpu_exits_with_message():
    r = _run("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_inflight_must_be_positive():
    r = _run("--inflight", "0")
    assert r.returncode == 2
    assert "--inflight" in r.stderr
