"""bench.py's command-line contract (no GPU needed): more than one GPU is
refused with a message and exit code 2 (every configuration fits one
MI355X; shots would shard one process per GPU), and the workloads the
documentation names are offered."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr


def test_workloads_are_offered():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("shot", "batch", "visco", "learned", "rtm2d", "gradient"):
        assert w in r.stdout
    assert "--batch-window" in r.stdout and "--imaging" in r.stdout
