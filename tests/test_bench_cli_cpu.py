"""bench.py runs on one GPU only: any other --gpus value exits with status 2
and a JSON line saying so, before it touches torch or a GPU (CPU)."""

import json
import os
import subprocess
import sys

from tests.conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    for n in (2, 8):
        r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)], capture_output=True,
                           text=True, timeout=120, cwd=ROOT)
        assert r.returncode == 2, r.stderr
        line = json.loads(r.stdout.strip().splitlines()[-1])
        assert line["n_gpus"] == n and "one GPU" in line["error"]
