"""bench.py's command line on the CPU: it refuses more than one GPU with a JSON
error line and exit status 2 (before touching the GPU), and lists its legs."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], capture_output=True, text=True,
                          timeout=120, cwd=ROOT)


def test_bench_refuses_several_gpus():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_bench_lists_its_legs():
    r = _bench("--help")
    assert r.returncode == 0
    for opt in ("--steps", "--warmup", "--reference-only", "--nufft-only", "--no-cpu", "--no-extra", "--volume"):
        assert opt in r.stdout, opt
