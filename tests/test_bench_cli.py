"""bench.py's command-line contract that needs no GPU: more than one GPU is
refused with a message (this build measures one), and the sample run's
profile pass only runs eagerly (graph replay is not traced)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), *args], cwd=ROOT, capture_output=True,
                          text=True, timeout=300)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_sample_profile_pass_needs_eager():
    r = _bench("--profile-pass", "--workload", "sample")
    assert r.returncode != 0
    assert "--eager" in r.stderr
