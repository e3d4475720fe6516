"""bench.py's command line without a GPU: more than one GPU is refused with
a message and exit status 2 before anything is imported or launched."""
import os
import subprocess
import sys


def test_bench_refuses_more_than_one_gpu():
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert 'exactly one GPU' in r.stderr
    assert r.stdout == ''
