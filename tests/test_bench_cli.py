"""bench.py's command line on the CPU: a request for more than one GPU exits with status 2 and
says why before anything touches a device (the measurement covers one GPU; states shard
across GPUs with no exchange)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr and '--gpus 2' in r.stderr
    assert r.stdout == ''
