"""bench.py's command line without a GPU: a request for more than one GPU
ends at once with exit status 2 and says why (the path is built for one GPU
per box), before anything touches a device; --help lists the workloads."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args],
                          capture_output=True, text=True, timeout=120, cwd=ROOT)


def test_more_than_one_gpu_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_help_lists_workloads():
    r = _bench("--help")
    assert r.returncode == 0
    for w in ("standard", "large", "small", "batch", "atoms"):
        assert w in r.stdout
