"""bench.py's command-line contract on a machine without a GPU: more than one
GPU is refused with a message and a non-zero status, before any device work."""
import subprocess
import sys

from helpers import ROOT


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""
