"""bench.py's refusals, run on the CPU: more than one GPU, and the largest
configuration (t-J 48 x 6) at an m whose blocks cannot fit one GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench(*args):
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + list(args), capture_output=True,
                          text=True, timeout=120)


def test_more_than_one_gpu_is_refused():
    r = _bench("--gpus", "2")
    assert r.returncode == 2 and "one GPU" in r.stderr


def test_c5_at_its_own_m_is_refused_with_the_reason():
    for extra in ([], ["--m", "8000"], ["--m", "4000"]):
        r = _bench("--workload", "c5", *extra)
        assert r.returncode == 2 and "does not fit one MI355X" in r.stderr, (extra, r.stderr)
