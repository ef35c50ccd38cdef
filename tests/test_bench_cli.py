"""CPU: bench.py measures one GPU; asked for more it exits with status 2 and
says so, before touching a GPU."""
import os
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "measures one GPU" in r.stderr and r.stdout == ""
