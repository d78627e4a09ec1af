"""Host-side behaviour of bench.py that needs no GPU: the hot path runs on one GPU, so a request for
more exits with status 2 and says so before touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1", "--warmup", "0"],
                         capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert out.returncode == 2
    assert "one GPU only" in out.stderr and "--gpus 2" in out.stderr
    assert out.stdout == ""
