"""bench.py measures one GPU: asked for several it exits with status 2 and a
message saying so, before it initialises a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr or "single MI355X" in p.stderr
    assert p.stdout.strip() == ""


def test_history_override_needs_a_single_cached_path():
    """--history rewrites the cached schedule of a profiler run; without
    --only and --schedule-cache it would have nothing to act on."""
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--history", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "--history" in p.stderr
    assert p.stdout.strip() == ""
