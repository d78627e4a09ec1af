"""Host-side behaviour that needs no GPU: bench.py runs the hot path on one GPU, so a request for
more exits with status 2 and says so before touching a device; and the solver library loader
raises, naming the path, when the built HIP library is missing (there is no CPU fallback)."""
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


def test_missing_library_fails_loudly():
    # the product path has no CPU fallback: without the built HIP library the loader raises
    env = dict(os.environ, DS_SMOKE_LIBRARY=os.path.join(ROOT, "no_such_dir", "libds_smoke.so"),
               PYTHONPATH=os.path.join(ROOT, "diff-solver_amd"))
    code = "from ds_smoke.solver import load_library\nload_library()\n"
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300, cwd=ROOT, env=env)
    assert out.returncode != 0
    assert "HIP smoke library not found" in out.stderr and "no_such_dir" in out.stderr
